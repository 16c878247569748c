"""The C-ABI library loads on any host and exports every declared symbol."""

import ctypes
import re

from conftest import ROOT


def declared():
    text = (ROOT / "include" / "wfdtd.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(wfd_\w+)\(", text, re.M)))


def test_library_exports_every_header_symbol(pkg):
    lib = pkg._lib.lib()
    names = declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
        assert name in pkg._lib.SIGNATURES, name


def test_version_and_error_text(pkg):
    lib = pkg._lib.lib()
    assert lib.wfd_version() == 1
    assert isinstance(lib.wfd_last_error(), bytes)


def test_bad_arguments_raise_value_error_without_a_device(pkg):
    """Argument checks run on the host before any launch."""
    import pytest

    with pytest.raises(ValueError):
        pkg._lib.call("wfd_advance", None, 1, None)
    plan = pkg._lib.YeePlan()
    with pytest.raises(ValueError, match="grid dimensions"):
        pkg._lib.call("wfd_yee_step", ctypes.byref(plan), None)


def test_inject_many_rejects_bad_tables_without_a_device(pkg):
    """wfd_adjoint_inject_many checks its plane table on the host: the plane
    count, null arrays and plane shapes, before any launch."""
    import pytest

    lib = pkg._lib
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every case fails first
    step = ctypes.c_void_p(0x2000)
    table = (lib.InjectSource * (lib.WFD_MAX_INJECT + 1))()
    with pytest.raises(ValueError, match="planes"):
        lib.call("wfd_adjoint_inject_many", ctypes.cast(table, ctypes.c_void_p),
                 lib.WFD_MAX_INJECT + 1, fake, fake, step, 8, 8, 8, 0, None)
    with pytest.raises(ValueError, match="arrays"):
        lib.call("wfd_adjoint_inject_many", ctypes.cast(table, ctypes.c_void_p), 1, fake,
                 fake, step, 8, 8, 8, 0, None)
    table[0] = lib.InjectSource(0x3000, 0x4000, 2, 1, 1, 3, 4, 13, 5)  # 13 % (3 * 4) != 0
    with pytest.raises(ValueError, match="shape"):
        lib.call("wfd_adjoint_inject_many", ctypes.cast(table, ctypes.c_void_p), 1, fake,
                 fake, step, 8, 8, 8, 0, None)
