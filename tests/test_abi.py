"""The C ABI from a plain C client (tests/abi_harness.c, gcc -std=c99
-Wall -Wextra -Werror against include/ugx.h):

- the header compiles as C and as C++;
- sizeof / offsetof of every field of ugx_options, ugx_result and
  ugx_problem equal those of the ctypes mirrors in ugacc_amd/ugx.py (a
  field added to one side only, or reordered, fails here);
- (GPU) a CCSD(T) through the C ABI reproduces a golden fixture, and the
  MAXITER failure path returns UGX_NOT_CONVERGED with the reference's text.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from tests import abi_util
from ugacc_amd import dump, ugx

ROOT = abi_util.ROOT


def _layout():
    out = subprocess.run([abi_util.harness(), "layout"], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles(lang, tmp_path):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "ugx.h"\nint main(void) { ugx_options o; ugx_default_options(&o); return o.maxiter; }\n')
    cc = "gcc" if lang == "c" else "g++"
    std = "-std=c99" if lang == "c" else "-std=c++11"
    subprocess.run([cc, std, "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)


@pytest.mark.parametrize("cname,pyt", [("ugx_options", ugx.Options), ("ugx_result", ugx.Result),
                                       ("ugx_problem", ugx.Problem)])
def test_struct_layout_matches_ctypes(cname, pyt):
    lay = _layout()[cname]
    assert lay["size"] == ctypes.sizeof(pyt)
    py = [(name, getattr(pyt, name).offset, getattr(pyt, name).size) for name, _ in pyt._fields_]
    c = [(name, off, size) for name, (off, size) in lay["fields"].items()]
    assert py == c


@pytest.mark.gpu
def test_solve_through_c_client_matches_fixture(tmp_path):
    from tests.gpu_util import announce
    announce("abi harness solve syn_o3v6_offdiag")
    d = dump.Dump(os.path.join(ROOT, "tests", "golden", "syn_o3v6_offdiag.ugxd"))
    fock, blocks = d.hamiltonian()
    (tmp_path / "sizes.txt").write_text("%d %d\n" % (d.sizes["no"], d.sizes["nv"]))
    np.ascontiguousarray(fock, dtype="<f8").tofile(str(tmp_path / "fock.bin"))
    for k, b in blocks.items():
        np.ascontiguousarray(b, dtype="<f8").tofile(str(tmp_path / (k + ".bin")))
    r = subprocess.run([abi_util.harness(), "solve", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    lines = [json.loads(l) for l in r.stdout.splitlines()]
    res, post, fail = lines[0], lines[1], lines[2]
    assert res["rc"] == ugx.OK and res["device_allocs"] == 0
    assert abs(res["e_mp2"] - d.scalars["e_mp2"]) < 1e-12
    assert abs(res["e_ccsd"] - d.scalars["eccsd"]) < 1e-10
    assert abs(res["e_t"] - d.scalars["e_t_tcorr"]) < 1e-10
    assert fail["rc_maxiter1"] == ugx.NOT_CONVERGED and fail["error"] == "Computation has not converged."
    # Lambda and the density energy through the C ABI: E_one + E_two = E_CCSD (f_ov = 0)
    assert post["rc_lambda"] == ugx.OK and post["rc_density"] == ugx.OK and post["niter_lambda"] > 0
    assert abs(post["e_one"] + post["e_two"] - res["e_ccsd"]) < 1e-9
