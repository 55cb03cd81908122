"""include/sun_abi.h must give gridT, physT and propT the binary layout of the
reference's own headers (main/grid.h, main/phys.h): same size, every member
at the same offset.  Both headers are compiled into small programs that print
offsetof() of every member; skipped where the reference tree is absent."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFMAIN = "/root/reference/main"

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REFMAIN, "phys.h")) or shutil.which("gcc") is None,
                                reason="reference headers not available")


def members(struct):
    text = open(os.path.join(ROOT, "include", "sun_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct _%s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        # drop the type: everything before the first declarator
        parts = decl.split(",")
        first = re.sub(r"^(const\s+)?(unsigned\s+)?[A-Za-z_][A-Za-z0-9_]*\s+", "", parts[0].strip())
        for p in [first] + parts[1:]:
            n = re.sub(r"[\*\s]|\[.*\]", "", p)
            names.append(n)
    return names


def offsets(include_lines, defines):
    prog = "#include <stdio.h>\n#include <stddef.h>\n" + include_lines + "\nint main(void){\n"
    for s in ("gridT", "physT", "propT"):
        prog += f'  printf("{s} size %zu\\n", sizeof({s}));\n'
        for m in members(s):
            prog += f'  printf("{s}.{m} %zu\\n", offsetof({s}, {m}));\n'
    prog += "  return 0;\n}\n"
    d = tempfile.mkdtemp()
    src, exe = os.path.join(d, "o.c"), os.path.join(d, "o")
    open(src, "w").write(prog)
    subprocess.run(["gcc", "-w", *defines, src, "-o", exe], check=True)
    return subprocess.run([exe], capture_output=True, text=True, check=True).stdout


def test_layout_identical_to_reference_headers():
    ours = offsets('#include "sun_abi.h"', [f"-I{ROOT}/include"])
    theirs = offsets('#include "grid.h"\n#include "phys.h"', [f"-I{REFMAIN}", "-DNOMPI", "-fcommon"])
    assert ours == theirs
    assert ours.count("\n") > 250
