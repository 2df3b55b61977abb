"""The C-ABI library loads and exports every symbol include/*.h declares."""
import glob
import os
import re

from tests.conftest import ROOT


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(kc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared():
    from ktcountermemorygreedy_amd import kc
    L = kc.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares signatures for exactly the header's set
    assert names == set(kc.EXPORTS), names ^ set(kc.EXPORTS)


def test_strerror_no_device_call():
    from ktcountermemorygreedy_amd import kc
    assert kc.lib().kc_strerror(kc.KC_EFULL) == b"table or FIRST list full"
