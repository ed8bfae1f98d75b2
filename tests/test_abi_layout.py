"""The ctypes mirrors of the C-ABI structs (m3tet_sph_amd/api.py, learned.py)
have the C layout of include/m3stokes.h and include/m3gnn.h: every field at
the C offset and the same struct size, checked against a program compiled
from the headers with the host C compiler."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pairs():
    from m3tet_sph_amd import api, learned
    return [("m3_mesh_desc", api.MeshDesc), ("m3_phys", api.Phys), ("m3_patera_opts", api.PateraOpts),
            ("m3_info", api.Info), ("m3_rock_params", api.RockParams), ("m3_diffusion_opts", api.DiffusionOpts),
            ("m3_diffusion_info", api.DiffusionInfo), ("m3_graph_desc", learned.GraphDesc)]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_ctypes_structs_match_the_headers(tmp_path):
    pairs = _pairs()
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "m3stokes.h"', '#include "m3gnn.h"',
             'int main(void) {', '  printf("{");']
    for k, (cname, cls) in enumerate(pairs):
        sep = "," if k else ""
        lines.append(f'  printf("{sep}\\"{cname}\\": [%zu", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'  printf(", %zu", offsetof({cname}, {f[0]}));')
        lines.append('  printf("]");')
    lines += ['  printf("}\\n");', '  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    for cname, cls in pairs:
        want = [__import__("ctypes").sizeof(cls)] + [getattr(cls, f[0]).offset for f in cls._fields_]
        assert got[cname] == want, (cname, got[cname], want)
