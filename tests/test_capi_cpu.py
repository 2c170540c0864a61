"""Host-side checks of the C ABI (no GPU needed): the library loads, exports
every function include/qm.h declares, and rejects bad arguments before it
touches the device."""
import ctypes as C
import os
import re

import pytest

from open_quadtree_mapping_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "qm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qm_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_table_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_default_params_are_reference_constants():
    p = _lib.default_params()
    assert p.n_hyp == 64 and p.idepth_min == 0.01 and p.idepth_max == 2.0
    assert abs(p.p1 - 0.003) < 1e-9 and abs(p.p2 - 0.01) < 1e-9
    assert list(p.bp_iters) == [4, 10, 10, 10] and p.lambda_ == 10.0
    assert p.min_refs == 11 and p.max_refs == 60


def test_invalid_arguments_fail_before_device_work():
    L = _lib.lib()
    rc = L.qm_quadtree(None, 8, 8, 8, None, None)
    assert rc == -1 and b"size" in L.qm_last_error()
    p = _lib.default_params(n_hyp=48)
    rc = L.qm_bp_extract(C.c_void_p(1), 64, 64, C.c_void_p(1), C.byref(p), C.c_void_p(1), C.c_void_p(1), None,
                         None)
    assert rc == -1 and b"n_hyp" in L.qm_last_error()
    p = _lib.default_params()
    rc = L.qm_upsample(None, None, 64, 64, 64, C.byref(p), None, None)
    assert rc == -1 and b"NULL" in L.qm_last_error()
    out = C.c_void_p()
    rc = L.qm_batch_create(None, 0, C.byref(p), C.byref(out))
    assert rc == -1 and out.value is None


def test_check_raises_with_message():
    with pytest.raises(_lib.QmError, match="size"):
        _lib.check(_lib.lib().qm_quadtree(None, 4, 4, 4, None, None))


def test_abi_version_matches_binding():
    assert _lib.lib().qm_abi_version() == _lib.ABI_VERSION
    hdr = open(os.path.join(ROOT, "include", "qm.h")).read()
    assert f"#define QM_ABI_VERSION {_lib.ABI_VERSION}" in hdr


def test_alternate_and_output_entry_points_reject_bad_arguments():
    L = _lib.lib()
    p = _lib.default_params()
    assert L.qm_quadtree_prior(None, 64, 48, None, None) == -1 and b"NULL" in L.qm_last_error()
    assert L.qm_prior_to_cost(None, 64, 48, None, C.byref(p), None, None, None) == -1
    assert b"NULL" in L.qm_last_error()
    bad = _lib.default_params(n_hyp=48)
    assert L.qm_prior_to_cost(C.c_void_p(1), 64, 48, C.c_void_p(1), C.byref(bad), C.c_void_p(1), None, None) == -1
    assert b"n_hyp" in L.qm_last_error()
    assert L.qm_upsample_local(None, None, 64, 48, 64, None, None) == -1 and b"NULL" in L.qm_last_error()
    assert L.qm_upsample_local(C.c_void_p(1), C.c_void_p(1), 64, 48, 32, C.c_void_p(1), None) == -1
    assert b"pitch" in L.qm_last_error()
    assert L.qm_sd_step(None, None, None, None, None, 64, 48, 64, None, None, None, 0, None, None, None) == -1
    assert L.qm_batch_set_tuning(None, 3, 1) == -1 and L.qm_batch_autotune(None, 1, None, None, None) == -1


def test_tuning_knob_names_match_header():
    hdr = open(os.path.join(ROOT, "include", "qm.h")).read()
    ids = dict((m.group(1).lower(), int(m.group(2))) for m in re.finditer(r"#define QM_TUNE_([A-Z_]+) (\d+)", hdr))
    assert ids == {k: i for i, k in enumerate(_lib.TUNE_KNOBS)}
    assert f"#define QM_NUM_TUNE {len(_lib.TUNE_KNOBS)}" in hdr


def test_fuse_noalloc_rejects_sizes_beyond_its_static_scratch():
    L = _lib.lib()
    one = C.c_void_p(1)
    rc = L.qm_fuse_noalloc(one, one, 2048, 1088, one, one, one, one, None, None)
    assert rc == -1 and b"2097152" in L.qm_last_error()
    rc = L.qm_fuse_noalloc(None, one, 752, 480, one, one, one, one, None, None)
    assert rc == -1 and b"NULL" in L.qm_last_error()


def test_release_scratch_without_work_is_a_no_op():
    assert _lib.lib().qm_release_scratch() == 0
