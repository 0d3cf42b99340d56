"""The allocation-free observable entry point against the allocating one.

qm_observable_expectation_noalloc_host runs the kernels of
qm_observable_expectation_host with its workspace in a fixed device array, so
the two return the same value bit for bit; both are also checked against the
oracle's term sum.  Two calls in a row, the second on a smaller state with fewer
terms, show that what the first call left in the array does not leak into the
second."""
from __future__ import annotations

import numpy as np
import pytest

from qulacs_amd import _lib
from tests.golden.gen import pauli_masks, random_pauli_terms, random_state
from tests.gpu_util import DevState, cmat, to_complex, u32

pytestmark = pytest.mark.gpu

MAX_TERMS = 128


def _tables(terms):
    masks = [pauli_masks(tg, pl) for tg, pl, _ in terms]
    k = len(terms)
    bfs = (_lib.C.c_ulonglong * k)(*[m[0] for m in masks])
    pfs = (_lib.C.c_ulonglong * k)(*[m[1] for m in masks])
    rots = u32([m[2] for m in masks])
    coefs = cmat([cf for *_, cf in terms])
    return bfs, pfs, rots, coefs


def _both(terms, s):
    bfs, pfs, rots, coefs = _tables(terms)
    a = to_complex(_lib.call("qm_observable_expectation_host", bfs, pfs, rots, coefs, len(terms), s.ptr, s.dim, None, 0))
    b = to_complex(_lib.call("qm_observable_expectation_noalloc_host", bfs, pfs, rots, coefs, len(terms), s.ptr, s.dim,
                             None, 0))
    return a, b


@pytest.mark.parametrize("sizes", [((20, 96), (12, 7)), ((16, MAX_TERMS), (5, 3)), ((23, 40), (10, 1))])
def test_noalloc_matches_allocating_entry_point(oracle, sizes):
    rng = np.random.default_rng(sum(n * 1000 + k for n, k in sizes))
    for n, k in sizes:  # the second call is smaller in qubits and terms
        print(f"noalloc observable: {n} qubits, {k} terms", flush=True)
        psi = random_state(rng, n)
        terms = random_pauli_terms(rng, n, k)
        with DevState.from_host(psi) as s:
            a, b = _both(terms, s)
        assert a.real == b.real and a.imag == b.imag, (n, k, a, b)
        want = sum(cf * oracle.expectation(tg, pl, psi) for tg, pl, cf in terms)
        scale = max(abs(want), sum(abs(cf) for *_, cf in terms))
        assert abs(b.real - want) <= 1e-10 * scale and abs(b.imag) <= 1e-10 * scale, (n, k, b, want)


def test_noalloc_z_only_and_flip_only_groups(oracle):
    rng = np.random.default_rng(7)
    n = 14
    psi = random_state(rng, n)
    zterms = [([0, 3, 9], [3, 3, 3], 0.5), ([13], [3], -1.25)]
    xterms = [([1, 2], [1, 2], 0.75), ([4, 13], [2, 1], 2.0)]
    with DevState.from_host(psi) as s:
        for terms in (zterms, xterms, zterms + xterms):
            a, b = _both(terms, s)
            assert a == b
            want = sum(cf * oracle.expectation(tg, pl, psi) for tg, pl, cf in terms)
            assert abs(b.real - want) <= 1e-12 * sum(abs(cf) for *_, cf in terms)


def test_noalloc_rejects_too_many_terms():
    rng = np.random.default_rng(3)
    n = 8
    terms = random_pauli_terms(rng, n, MAX_TERMS + 1)
    bfs, pfs, rots, coefs = _tables(terms)
    with DevState.from_host(random_state(rng, n)) as s:
        with pytest.raises(Exception, match="exceed"):
            _lib.call("qm_observable_expectation_noalloc_host", bfs, pfs, rots, coefs, len(terms), s.ptr, s.dim,
                      None, 0)
        # the error leaves the entry point usable
        a, b = _both(terms[:5], s)
        assert a == b
