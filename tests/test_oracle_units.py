"""Unit pins of the CPU oracle (CPU only): the random-number generator, the
LTC table blob, the edge trees and the library's exported symbols."""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

from delta_ray_amd import _abi
from tests import helpers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# Philox4x32-10 known-answer vectors published with the generator (Random123
# distribution, kat_vectors: "philox4x32 10 <ctr[4]> <key[2]> <out[4]>").
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    lib = oracle.load()
    k = (ctypes.c_uint32 * 2)(*key)
    c = (ctypes.c_uint32 * 4)(*ctr)
    out = (ctypes.c_uint32 * 4)()
    lib.or_philox4x32(k, c, out)
    assert tuple(out) == expect
