"""C-ABI checks that need no GPU: the library loads, exports every symbol the
headers declare, validates geometry and reports errors without exiting."""
import ctypes
import re
from dataclasses import replace
from pathlib import Path

import pytest

from gpu_amd import core
from gpu_amd.geometry import Geometry, adjoint_64, multisource_256, plain_sart_128, repo_scale_512

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> list[str]:
    """Names of the function prototypes in include/<header>."""
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    text = text.replace('extern "C" {', "")
    names = []
    for stmt in " ".join(text.split()).split(";"):
        stmt = stmt.strip().lstrip("}").strip()
        if not stmt or stmt.startswith("typedef"):
            continue
        m = re.match(r"^(?:const\s+)?[A-Za-z_]\w*[\s\*]+(\w+)\s*\(", stmt)
        if m:
            names.append(m.group(1))
    return sorted(set(names))


def test_all_declared_symbols_exported():
    lib = ctypes.CDLL(str(core.LIB_PATH))
    for header in ("ct_core.h", "ct_compat.h"):
        names = declared(header)
        assert names, header
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, f"{header}: {missing}"


def test_abi_version():
    assert core.lib().ct_abi_version() == 1


@pytest.mark.parametrize("mk", [plain_sart_128, repo_scale_512, adjoint_64, multisource_256])
def test_workload_geometries_valid(mk):
    assert core.validate_geometry(mk()) == 0


@pytest.mark.parametrize("bad,code", [
    (dict(nx=0), 2), (dict(vx=-1.0), 2), (dict(du=0.0), 2), (dict(nviews=0), 2),
    (dict(dso=-600.0), 2), (dict(dso=50.0), 2), (dict(nsources=0), 2),
    (dict(nx=2048, ny=2048, nz=1024, vx=0.1), 6),
])
def test_invalid_geometry_rejected(bad, code):
    g = replace(plain_sart_128(), **bad)
    assert core.validate_geometry(g) == code


def test_view_angles_follow_reference_sign():
    g = plain_sart_128()
    s, c = core.view_angle(g, 45)           # theta = 90 deg, iterative operators use -theta
    assert s == pytest.approx(-1.0, abs=1e-6) and c == pytest.approx(0.0, abs=1e-6)
    s, c = core.view_angle(g, 45, fdk=True)
    assert s == pytest.approx(1.0, abs=1e-6)
    with pytest.raises(core.CtError):
        core.view_angle(g, 180)


def test_no_device_reports_error_not_exit():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(core.CtError) as e:
        core.Context(0)
    assert e.value.status == 8   # CT_ERR_DEVICE


def test_status_strings():
    lib = core.lib()
    for s in range(10):
        assert lib.ct_status_string(s)
