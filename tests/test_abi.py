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
    assert kc.lib().kc_strerror(kc.KC_EFULL) == b"table full"


# ctypes mirror of each C struct of the headers (layout must match exactly)
STRUCTS = {
    "kc_table_config": "TableConfig", "kc_stats": "Stats", "kc_tune": "Tune", "kc_plan": "Plan",
    "kc_batch_info": "BatchInfo", "kc_ingest_stats": "IngestStats", "kc_pass_plan": "PassPlan",
    "kc_genome_spec": "GenomeSpec", "kc_read_spec": "ReadSpec", "kc_request_result": "RequestResult",
}


def test_struct_layouts_match_headers(tmp_path):
    """sizeof and every field offset of the C structs, compiled from the
    headers with the host compiler, equal the binding's ctypes structs."""
    import ctypes
    import subprocess

    from ktcountermemorygreedy_amd import kc
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "kc_api.h"', '#include "kc_host.h"',
             '#include "kc_server.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        cls = getattr(kc, pyname)
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    got = {}
    for line in out:
        a, b, c = line.split()
        got[(a, b)] = int(c)
    for cname, pyname in STRUCTS.items():
        cls = getattr(kc, pyname)
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)
