"""The C-ABI boundary: the library loads, exports every declared symbol,
and its table/struct layouts are the reference's (src/ops.h:43-152,
app/app_ccs.h:20-24, app/app_lapack.h:17-20).  No device calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

from gcge_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for h in sorted(os.listdir(INCLUDE)):
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        # function declarations at file scope (the struct members are
        # function pointers, written "(*name)", and are not matched)
        for m in re.finditer(r"^\s*(?:void|int|double|int64_t)\s*\*?\s*(\w+)\s*\(", text,
                             re.M):
            names.append(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table declares each of them
    assert set(names) <= set(abi.EXPORTS), set(names) - set(abi.EXPORTS)


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include HEADER
#define F(x) printf("%s %zu\n", #x, offsetof(struct OPS_, x));
int main(void) {
  FIELDS
  printf("sizeof %zu\n", sizeof(struct OPS_));
#ifdef CCS_T
  printf("ccs %zu %zu %zu %zu %zu %zu\n", sizeof(CCS_T), offsetof(CCS_T, data),
         offsetof(CCS_T, i_row), offsetof(CCS_T, j_col), offsetof(CCS_T, nrows),
         offsetof(CCS_T, ncols));
  printf("dense %zu %zu %zu %zu %zu\n", sizeof(DENSE_T), offsetof(DENSE_T, data),
         offsetof(DENSE_T, nrows), offsetof(DENSE_T, ncols), offsetof(DENSE_T, ldd));
#endif
  return 0;
}
"""


def run_probe(tmp_path, header, extra, ccs_t, dense_t, tag):
    fields = " ".join(f"F({n})" for n, _ in abi.OPS_FIELDS)
    src = tmp_path / f"probe_{tag}.c"
    src.write_text(PROBE.replace("FIELDS", fields))
    exe = tmp_path / f"probe_{tag}"
    cmd = ["gcc", "-std=gnu99", f"-DHEADER={header}", f"-DCCS_T={ccs_t}",
           f"-DDENSE_T={dense_t}"] + extra + [str(src), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return out.split("\n")


def test_layout_matches_ctypes(tmp_path):
    lines = run_probe(tmp_path, '"gcge_ops.h"', [f"-I{INCLUDE}"], "gcge_ccs_mat",
                      "gcge_dense_mv", "ours")
    offs = dict(l.split()[:2] for l in lines if l and not l.startswith(("ccs", "dense")))
    for name, _ in abi.OPS_FIELDS:
        assert int(offs[name]) == getattr(abi.OPS, name).offset, name
    assert int(offs["sizeof"]) == C.sizeof(abi.OPS)
    ccs = [int(x) for x in next(l for l in lines if l.startswith("ccs")).split()[1:]]
    assert ccs == [C.sizeof(abi.CCSMat)] + [getattr(abi.CCSMat, f).offset for f in
                                            ("data", "i_row", "j_col", "nrows", "ncols")]
    dn = [int(x) for x in next(l for l in lines if l.startswith("dense")).split()[1:]]
    assert dn == [C.sizeof(abi.DenseMV)] + [getattr(abi.DenseMV, f).offset for f in
                                            ("data", "nrows", "ncols", "ldd")]


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"),
                    reason="reference sources not present")
def test_layout_matches_reference_header(tmp_path):
    ref = "/root/reference"
    extra = [f"-I{ref}/src", f"-I{ref}/app", "-include", os.path.join(ROOT, "oracle",
                                                                      "ref_config.h")]
    theirs = run_probe(tmp_path, '"app_ccs.h"', extra, "CCSMAT", "LAPACKVEC", "ref")
    ours = run_probe(tmp_path, '"gcge_ops.h"', [f"-I{INCLUDE}"], "gcge_ccs_mat",
                     "gcge_dense_mv", "ours2")
    assert theirs == ours


SERVICE_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "gcge_hip.h"
#define R(x) printf("req %s %zu\n", #x, offsetof(gcge_request, x));
#define S(x) printf("res %s %zu\n", #x, offsetof(gcge_result, x));
#define E(x) printf("eps %s %zu\n", #x, offsetof(gcge_eps_options, x));
#define G(x) printf("grid %s %zu\n", #x, offsetof(gcge_grid_op, x));
int main(void) {
  FIELDS
  printf("req sizeof %zu\nres sizeof %zu\neps sizeof %zu\ngrid sizeof %zu\n",
         sizeof(gcge_request), sizeof(gcge_result), sizeof(gcge_eps_options),
         sizeof(gcge_grid_op));
  return 0;
}
"""


def test_service_structs_match_ctypes(tmp_path):
    fields = " ".join([f"R({n})" for n, _ in abi.Request._fields_] +
                      [f"S({n})" for n, _ in abi.Result._fields_] +
                      [f"E({n})" for n, _ in abi.EpsOptions._fields_] +
                      [f"G({n})" for n, _ in abi.GridOp._fields_])
    src = tmp_path / "svc.c"
    src.write_text(SERVICE_PROBE.replace("FIELDS", fields))
    exe = tmp_path / "svc"
    subprocess.run(["gcc", "-std=gnu99", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = {(out[i], out[i + 1]): int(out[i + 2]) for i in range(0, len(out), 3)}
    for tag, cls in (("req", abi.Request), ("res", abi.Result), ("eps", abi.EpsOptions),
                     ("grid", abi.GridOp)):
        for n, _ in cls._fields_:
            assert got[(tag, n)] == getattr(cls, n).offset, (tag, n)
        assert got[(tag, "sizeof")] == C.sizeof(cls)
