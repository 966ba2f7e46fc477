"""C-ABI surface: libpict_hip.so loads without a GPU, exports every function include/pict_hip.h declares, and the
ctypes mirror (pict_amd/_abi.py) has the same struct layouts as the C header (checked with a gcc-compiled probe)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from pict_amd import _abi

HEADER = _abi.HEADER_PATH


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pict_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    declared = _declared_functions()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared) == set(_abi.SIGNATURES), set(declared) ^ set(_abi.SIGNATURES)
    assert lib.pict_abi_version() == _abi.ABI_VERSION == 2


def test_struct_layout_matches_header():
    structs = {
        "pict_boundary_desc": _abi.BoundaryDesc,
        "pict_block_desc": _abi.BlockDesc,
        "pict_domain_desc": _abi.DomainDesc,
        "pict_csr_desc": _abi.CsrDesc,
        "pict_solve_opts": _abi.SolveOpts,
        "pict_solve_info": _abi.SolveInfo,
        "pict_step_opts": _abi.StepOpts,
        "pict_step_tape": _abi.StepTape,
        "pict_batch_config": _abi.BatchConfig,
        "pict_batch_info": _abi.BatchInfo,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for cname, py in structs.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "probe.c")
        exe = os.path.join(tmp, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c99", src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln.strip():
            parts = ln.split()
            got[(parts[0], parts[1])] = int(parts[-1])
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        _abi.load_library(str(tmp_path / "libnothere.so"))


def test_stage_functions_refuse_cpu_domains():
    """There is no CPU fallback: a stage call on a host-memory domain raises instead of computing."""
    import torch
    import scenarios
    from pict_amd import PISOtorch as P
    d = scenarios.cavity(dtype=torch.float64, device="cpu")
    with pytest.raises(RuntimeError):
        P.SetupAdvectionMatrix(d, torch.tensor([0.1], dtype=torch.float64), 0)
    with pytest.raises(RuntimeError):
        P.Domain(2, torch.tensor([0.1]), "x", torch.float32, "cpu")
