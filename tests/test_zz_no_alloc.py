"""The allocation-free canonical export (mdx_pairs_canonical_noalloc)
against the allocating one (mdx_pairs_canonical), through the C ABI.

Bar: bit-exact (integer triples).  Two exports in a row share the device
staging array, the second from a smaller system, and each must equal its own
system's allocating export; the count-only, truncated and over-limit forms
follow mdx_pairs_canonical's contract or fail with MDX_ERR_CAPACITY.
"""
import ctypes

import numpy as np
import pytest

from mdacp_amd import workloads as W
from mdacp_amd.mdx import MdxError, System

pytestmark = pytest.mark.gpu

NOALLOC_PAIRS = 1 << 20   # MDX_NOALLOC_PAIRS (include/mdx.h)
MDX_ERR_CAPACITY = -2


def noalloc(sys_, cap=None):
    """(rc, count, triples) of mdx_pairs_canonical_noalloc; cap None = count only."""
    count = ctypes.c_int64(-1)
    if cap is None:
        rc = sys_.lib.mdx_pairs_canonical_noalloc(sys_.h, None, 0, ctypes.byref(count))
        return rc, count.value, None
    out = np.full((max(cap, 1), 3), -7, dtype=np.int32)
    rc = sys_.lib.mdx_pairs_canonical_noalloc(
        sys_.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap, ctypes.byref(count))
    return rc, count.value, out[:cap]


def system(state):
    g = System.from_state(state)
    g.make_pairlist()
    return g


def test_noalloc_export_equals_allocating_export_twice_second_smaller():
    big = W.perturbed(W.plain_small(), 0.25, 7)                    # N = 4000
    small = W.perturbed(W.fcc_state(5, 0.7, 1.0, 11), 0.2, 12)     # N = 500
    with system(big) as a, system(small) as b:
        for g, label in ((a, "N=4000"), (b, "N=500")):
            print(f"noalloc export {label}", flush=True)
            ref = g.pairs()
            rc, cnt, tri = noalloc(g, len(ref))
            assert rc == 0, g.lib.mdx_last_error(g.h).decode()
            assert cnt == len(ref)
            np.testing.assert_array_equal(tri, ref)
        # the first system again, after the smaller one overwrote the staging
        rc, cnt, tri = noalloc(a, cnt_a := a.pair_number())
        assert rc == 0 and cnt == cnt_a
        np.testing.assert_array_equal(tri, a.pairs())


def test_noalloc_count_only_and_truncated():
    s = W.perturbed(W.fcc_state(6, 0.5, 1.0, 21), 0.3, 22)
    with system(s) as g:
        ref = g.pairs()
        rc, cnt, _ = noalloc(g)
        assert rc == 0 and cnt == len(ref)
        k = len(ref) // 3
        rc, cnt, tri = noalloc(g, k)
        assert rc == 0 and cnt == len(ref)
        np.testing.assert_array_equal(tri, ref[:k])


def test_noalloc_over_limit_is_refused_and_the_system_stays_usable():
    s = W.fcc_state(30, 0.5, 1.0, 23)   # N = 108,000: 4,212,000 pairs
    with system(s) as g:
        rc, cnt, _ = noalloc(g, 16)
        assert rc == MDX_ERR_CAPACITY
        assert cnt == 39 * s.n > NOALLOC_PAIRS
        assert "allocation-free" in g.lib.mdx_last_error(g.h).decode()
        assert g.pair_number() == 39 * s.n
    small = W.fcc_state(5, 0.5, 1.0, 24)
    with system(small) as g:
        rc, cnt, tri = noalloc(g, 39 * small.n)
        assert rc == 0 and cnt == 39 * small.n
        np.testing.assert_array_equal(tri, g.pairs())


def test_noalloc_without_pairlist_is_a_state_error():
    with System.from_state(W.fcc_state(5, 0.5, 1.0, 25)) as g:
        with pytest.raises(MdxError):
            g._chk(noalloc(g)[0])
