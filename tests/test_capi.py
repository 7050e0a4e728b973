"""The C ABI library loads and exports every entry point include/*.h
declares; the Python binding covers them all; the generators produce the
workloads they describe. CPU only (no compute calls)."""
import ctypes as C
import glob
import os
import re

import numpy as np

from sovereign_amd import _capi, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(sov_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) > 20
    lib = C.CDLL(_capi.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert declared_functions() == set(_capi.SIGNATURES)
    _capi.load()


def test_random_lp_generator_shape():
    lp = P.random_sparse_lp(512, 2048, 6, seed=1)
    assert lp.col_ptr[-1] == 6 * 2048
    for j in range(0, 2048, 97):
        rows = lp.row_idx[lp.col_ptr[j]:lp.col_ptr[j + 1]]
        assert len(set(rows.tolist())) == 6 and np.all(np.diff(rows) > 0)
    A = lp.as_model().dense()
    np.testing.assert_allclose(A @ lp.x0, lp.b, atol=1e-12)
    np.testing.assert_allclose(A.T @ lp.y0 + lp.s0, lp.c, atol=1e-12)
    assert lp.x0.min() >= 0.1 and lp.s0.min() >= 0.1


def test_transport_generator_balanced():
    m = P.transport_balanced(50, 50)
    assert (m.n_vars, m.n_rows, m.nnz) == (2500, 100, 5000)
    assert abs(m.rhs[:50].sum() - m.rhs[50:].sum()) < 1e-9
    assert m.obj.min() >= 1 and m.obj.max() <= 20
