"""The edge-tree traversal stack of the selection kernel: its top entries in
LDS and older ones moved to private memory one at a time must behave exactly
like a plain array stack (no reference counterpart; a self-check through
dr_test_tree_stack, which drives both with the same random pushes and pops on
every lane of 64 workgroups, deep enough to spill and reload; half the waves
carry the alternative pmfs of the learned mixture, half do not)."""
import ctypes

import pytest

from delta_ray_amd import _abi

pytestmark = pytest.mark.gpu


def test_lds_stack_matches_array_stack(gpu):
    lib = _abi.load_library()
    out = (ctypes.c_int64 * 2)()
    _abi.check(lib, lib.dr_test_tree_stack(4000, 17, out, None), "dr_test_tree_stack")
    assert out[0] == 0, "%d pops differ" % out[0]
    # the walk reaches the full depth (32), far past the 10 LDS entries
    assert out[1] == 32, out[1]
