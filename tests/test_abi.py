"""CPU checks of the C ABI: both libraries load (no GPU needed), export every
function their headers declare, and the kernel table covers every id."""
import ctypes as C
import re
from pathlib import Path

import pytest

import kutta_amd

ROOT = Path(__file__).resolve().parent.parent


def declared_functions(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(k[ha]_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,loader", [("kutta_hip.h", kutta_amd.hip_lib),
                                           ("kutta_amd.h", kutta_amd.host_lib)])
def test_library_exports_every_declared_symbol(header, loader):
    lib = loader()
    names = declared_functions(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_kernel_ids_have_names_and_match_python_enum():
    lib = kutta_amd.hip_lib()
    for k in kutta_amd.Kernel:
        name = lib.kh_kernel_name(int(k))
        assert name is not None, k
        assert name.decode() == k.name
    assert lib.kh_kernel_name(0) is None
    assert lib.kh_kernel_name(9999) is None


def test_kernel_enum_matches_header():
    text = (ROOT / "include" / "kutta_hip.h").read_text()
    body = text[text.index("typedef enum kh_kernel"):text.index("} kh_kernel;")]
    ids, value = {}, -1
    for m in re.finditer(r"^\s*(KH_\w+)(?:\s*=\s*(\d+))?,", body, flags=re.M):
        value = int(m.group(2)) if m.group(2) else value + 1
        ids[m.group(1)] = value
    ids.pop("KH_KERNEL_NONE")
    ids.pop("KH_KERNEL_COUNT_MARK", None)
    for k in kutta_amd.Kernel:
        assert ids["KH_" + k.name.upper()] == int(k), k
    assert sorted(ids) == sorted("KH_" + k.name.upper() for k in kutta_amd.Kernel)


def test_last_error_is_a_string_without_a_device():
    assert isinstance(kutta_amd.hip_lib().kh_last_error(), bytes)
