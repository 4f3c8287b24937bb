"""sphfem_tool_create refuses a tool description with an index outside its array (element nodes, CSR
pattern, segment marks, fixed-boundary elements), a non-finite node, a non-positive lumped mass, a
missing required array or empty sizes -- before anything is copied, so neither the host set-up nor a
device kernel can index out of bounds.  The check runs before the context is used: the test passes no
context and needs no GPU."""
import ctypes as C

import numpy as np
import pytest

from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

# order of the arrays solver.tool_desc keeps alive
POS, ELEM, LAYER, ROWPTR, COLIND, K, M, F, T, LEFT, RIGHT, N, MARKS, NFLAGS, FIXED, FFLAGS = range(16)


def _corruptions(nn, ne, nz):
    return [
        ("element node id = node_num", ELEM, lambda a: a.__setitem__((3, 1), nn)),
        ("element node id -1", ELEM, lambda a: a.__setitem__((0, 0), -1)),
        ("element colour -1", ELEM, lambda a: a.__setitem__((5, 3), -1)),
        ("CSR column = node_num", COLIND, lambda a: a.__setitem__(7, nn)),
        ("CSR rows start at 1", ROWPTR, lambda a: a.__setitem__(0, 1)),
        ("CSR rows end past nz", ROWPTR, lambda a: a.__setitem__(nn, nz + 1)),
        ("CSR rows decrease", ROWPTR, lambda a: a.__setitem__(2, a[3] + 1)),
        ("segment left node = node_num", MARKS, lambda a: a.__setitem__((1, 2), nn)),
        ("segment element = ele_num", MARKS, lambda a: a.__setitem__((0, 0), ne)),
        ("fixed element = ele_num", FIXED, lambda a: a.__setitem__(0, ne)),
        ("node at NaN", POS, lambda a: a.__setitem__((4, 0), np.nan)),
        ("zero lumped mass", M, lambda a: a.__setitem__(2, 0.0)),
    ]


def test_invalid_tool_descriptions_are_refused():
    case = cases.small_cutting_case(nx=32, ny=12)
    lib = _lib.lib()
    d, keep = S.tool_desc(case)
    nn, ne, nz = d.node_num, d.ele_num, d.nz_num
    h = C.c_void_p()
    for label, which, corrupt in _corruptions(nn, ne, nz):
        saved = keep[which].copy()
        corrupt(keep[which])
        assert lib.sphfem_tool_create(None, C.byref(d), C.byref(h)) == _lib.ERR_ARG, label
        assert not h.value, label
        keep[which][...] = saved
    for field in ("K", "marks", "csr_col_ind"):
        saved = getattr(d, field)
        setattr(d, field, None)
        assert lib.sphfem_tool_create(None, C.byref(d), C.byref(h)) == _lib.ERR_ARG, field
        setattr(d, field, saved)
    for field in ("node_num", "ele_num", "nz_num", "seg_num"):
        saved = getattr(d, field)
        setattr(d, field, 0)
        assert lib.sphfem_tool_create(None, C.byref(d), C.byref(h)) == _lib.ERR_ARG, field
        setattr(d, field, saved)
    d.fixed_num = -1
    assert lib.sphfem_tool_create(None, C.byref(d), C.byref(h)) == _lib.ERR_ARG
