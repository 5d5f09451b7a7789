"""C ABI checks that need no GPU: the library loads, exports every function
include/nbx.h declares, keeps the reference constructor defaults, and rejects
invalid options before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

import nbodylib_amd as nb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "nbx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header():
    lib = nb.lib()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(nb.EXPORTS) <= set(names)


def test_params_defaults_match_reference():
    """KDTree.h:229-245: bucket 16, TPHYS, KEPAN, kernres 1000, spread split,
    aniso 0, no scaling, no period."""
    p = nb.Params()
    nb.lib().nbx_params_default(ctypes.byref(p))
    assert (p.bucket_size, p.treetype, p.kerntype, p.kernres) == (16, nb.TPHYS, nb.KEPAN, 1000)
    assert (p.splitting, p.aniso, p.scalespace, p.periodic) == (0, 0, 0, 0)
    assert p.form == nb.FORM_PLAIN


def _create(p, n=10, pos=True, vel=False):
    import numpy as np

    a = np.zeros(3 * n)
    h = ctypes.c_void_p()
    ptr = ctypes.c_void_p(a.ctypes.data) if pos else None
    vptr = ctypes.c_void_p(a.ctypes.data) if vel else None
    return nb.lib().nbx_tree_create(ctypes.byref(p), ptr, vptr, None, n, ctypes.byref(h))


@pytest.mark.parametrize("field,value,code", [
    ("treetype", 7, nb.NBX_ERR_UNSUPPORTED),   # KDTree.cxx:1096-1099 "Error in type of tree specified"
    ("treetype", nb.TPROJ, nb.NBX_ERR_UNSUPPORTED),
    ("splitting", 1, nb.NBX_ERR_UNSUPPORTED),
    ("scalespace", 1, nb.NBX_ERR_UNSUPPORTED),
])
def test_invalid_options_rejected(field, value, code):
    p = nb.params()
    setattr(p, field, value)
    assert _create(p) == code
    assert len(nb.lib().nbx_last_error()) > 0


def test_phase_tree_needs_velocities_and_aniso():
    p = nb.params(treetype=nb.TPHS, aniso=-1)
    assert _create(p, vel=False) == nb.NBX_ERR_ARG
    p = nb.params(treetype=nb.TPHS, aniso=0)
    assert _create(p, vel=True) == nb.NBX_ERR_UNSUPPORTED


def test_velocity_tree_needs_velocities_and_no_period():
    """TVEL builds on the velocities; a periodic TVEL tree is rejected (the
    reference would apply the period to velocity images)."""
    p = nb.params(treetype=nb.TVEL)
    assert _create(p, vel=False) == nb.NBX_ERR_ARG
    p = nb.params(treetype=nb.TVEL, period=1.0)
    assert _create(p, vel=True) == nb.NBX_ERR_UNSUPPORTED


def test_null_and_size_arguments():
    p = nb.params()
    assert _create(p, pos=False) == nb.NBX_ERR_ARG
    assert _create(p, n=0) == nb.NBX_ERR_ARG
    bad = nb.params(period=0.0)
    assert _create(bad) == nb.NBX_ERR_ARG


def test_python_mirror_fails_loudly_without_library(tmp_path, monkeypatch):
    import importlib

    monkeypatch.setattr(nb, "LIBPATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(nb, "_lib", None)
    with pytest.raises(ImportError):
        nb.lib()
    monkeypatch.undo()
    importlib.reload(nb)
    nb.lib()


def test_fof_lists_from_labels():
    """nbx_fof_lists (host only): each group's members linked once, in slot
    order, from its head; tail of the head is the last member; len[g] is the
    group size; particles in no group are lists of one."""
    import oracle
    from nbodylib_amd import datagen

    n = 3000
    pos = datagen.uniform(n, 1.0, 53)
    lab, ng = oracle.fof_brute(pos, 0.6 * datagen.mean_spacing(n, 1.0), 3)
    assert ng > 5
    order = np.random.default_rng(53).permutation(n).astype(np.int32)  # slot -> ID
    head, nxt, tail, ln = nb.fof_lists(lab, order, ng)
    g_of_slot = lab[order]
    for g in range(1, ng + 1):
        slots = np.nonzero(g_of_slot == g)[0]
        assert ln[g] == slots.size
        walk, s = [], slots[0]
        while s != -1:
            walk.append(s)
            s = nxt[s]
        assert walk == slots.tolist()
        assert np.all(head[slots] == slots[0]) and tail[slots[0]] == slots[-1]
    lone = np.nonzero(g_of_slot == 0)[0]
    assert np.all(head[lone] == lone) and np.all(tail[lone] == lone) and np.all(nxt[lone] == -1)
    assert ln[0] == 0 and not ln[ng + 1:].any()
    with pytest.raises(nb.NbxError):
        nb.fof_lists(lab, order[::-1] * 0 + n, ng)
