"""The C ABI: libxba.so loads (no GPU needed to load it) and exports every entry point include/xba.h declares;
the host-side packing of Problems into the flat batch layout is a bijection on landmarks and observations."""
import ctypes
import os
import re

import numpy as np

from xclean_amd import _lib, synth
from xclean_amd.window import _Packed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "xba.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(xba_\w+)\s*\(", src, re.M)))


def test_header_declares_the_expected_entry_points():
    assert set(_declared()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.xba_version


def test_strerror_and_version_need_no_gpu():
    L = _lib.load()
    assert L.xba_version().decode().startswith("xba")
    assert L.xba_strerror(-3).decode().startswith("a landmark has more")


def test_packing_sorts_landmarks_by_slot_range_and_round_trips():
    pb, _ = synth.make_window(P=7, L=300, views=(1, 7), seed=[4, 4], posefree=True, stereo="prior")
    w = _Packed(pb)
    owner = np.repeat(np.arange(w.L), np.diff(w.ob_start))
    lo, hi = w.lm_host.astype(np.int64).copy(), w.lm_host.astype(np.int64).copy()
    np.minimum.at(lo, owner, w.ob_pose)
    np.maximum.at(hi, owner, w.ob_pose)
    key = lo * 1000 + hi              # narrow landmarks only here (P = 7): ordered by (first slot, last slot)
    assert np.all(np.diff(key) >= 0)
    assert w.L == pb.L + len(pb.pf.host) and w.M == pb.M + len(pb.pf.host)
    lm_all = np.concatenate([pb.lmp, pb.pf.lmp])
    np.testing.assert_array_equal(w.unpermute_lm(w.lmp), lm_all)
    z_all = np.concatenate([pb.ob_z, pb.pf.z])
    np.testing.assert_array_equal(w.unpermute_obs(w.ob_z), z_all)
    # every observation still belongs to its landmark
    owner = np.repeat(np.arange(w.L), np.diff(w.ob_start))
    orig_owner = np.concatenate([np.repeat(np.arange(pb.L), np.diff(pb.ob_start)), pb.L + np.arange(len(pb.pf.host))])
    np.testing.assert_array_equal(w.perm[owner], orig_owner[w.obs_src])
