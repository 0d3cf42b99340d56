"""Writes tests/golden/*.npz: input/output vectors produced by the reference csim.

The reference csim is compiled from its own sources by `make -C oracle ref`
into oracle/_ref/libcsim_ref.so; this script drives it through
oracle/ref_bridge.cpp on small seeded inputs and stores inputs and outputs.
The fixtures are data only; tests compare the in-repo oracle and the HIP
library against them without needing the reference on the machine.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402
from tests.golden.gen import haar_unitary, random_pauli_terms, random_state  # noqa: E402

OUT = Path(__file__).resolve().parent


def main() -> None:
    O.build(ref=True)
    ref = O.Reference()
    ref.set_num_threads(1)
    rng = np.random.default_rng(20261202)
    n = 7
    dim = 1 << n

    # ------------------------------------------------------------------ state-vector gates
    cases = []  # (name, kind, targets, controls, cvals, param, matrix)
    for t in range(n):
        cases.append(("dense1", "dense", [t], [], [], 0.0, haar_unitary(rng, 1)))
    for (a, b) in [(0, 1), (1, 0), (0, 6), (6, 2), (3, 5), (5, 4)]:
        cases.append(("dense2", "dense", [a, b], [], [], 0.0, haar_unitary(rng, 2)))
    for tg in ([0, 1, 2], [6, 0, 3], [2, 5, 1, 4], [0, 1, 2, 3, 4], [6, 1, 3, 0, 5], [4, 2, 0, 6, 1, 3]):
        cases.append((f"dense{len(tg)}", "dense", tg, [], [], 0.0, haar_unitary(rng, len(tg))))
    for (c, cv, t) in [(0, 1, 1), (5, 0, 2), (2, 1, 6)]:
        cases.append(("cdense1", "cdense", [t], [c], [cv], 0.0, haar_unitary(rng, 1)))
    cases.append(("cdense_multi", "cdense", [3], [0, 5], [1, 0], 0.0, haar_unitary(rng, 1)))
    cases.append(("c1dense2", "cdense", [1, 4], [6], [1], 0.0, haar_unitary(rng, 2)))
    cases.append(("ccdense3", "cdense", [0, 2, 5], [1, 6], [0, 1], 0.0, haar_unitary(rng, 3)))
    for nm in ("X", "Y", "Z", "H", "S", "Sdag", "T", "Tdag", "sqrtX", "sqrtXdag", "sqrtY", "sqrtYdag", "P0", "P1"):
        for t in (0, 3, 6):
            cases.append((nm, "named", [t], [], [], 0.0, None))
    for nm in ("RX", "RY", "RZ"):
        for t in (0, 4, 6):
            cases.append((nm, "rot", [t], [], [], float(rng.uniform(0, 2 * np.pi)), None))
    for nm in ("CNOT", "CZ", "SWAP"):
        for (a, b) in [(0, 1), (1, 0), (0, 6), (6, 3), (2, 5)]:
            cases.append((nm, "two", [b], [a], [], 0.0, None))
    for t in (0, 5):
        d = np.exp(1j * rng.uniform(0, 2 * np.pi, 2))
        cases.append(("diag1", "diag1", [t], [], [], 0.0, d))
        cases.append(("phase1", "phase1", [t], [], [], 0.0, np.exp(1j * rng.uniform(0, 2 * np.pi, 1))))
    for tg, cl, cv in (([0, 3], [], []), ([5, 1, 2], [], []), ([2], [4], [1]), ([6, 0], [3], [0])):
        d = np.exp(1j * rng.uniform(0, 2 * np.pi, 1 << len(tg)))
        cases.append(("cdiag", "cdiag", tg, cl, cv, 0.0, d))
    pauli_sets = []
    for _ in range(8):
        k = int(rng.integers(1, n + 1))
        tg = rng.choice(n, size=k, replace=False)
        pl = rng.integers(0, 4, size=k)
        pauli_sets.append((tg.tolist(), pl.tolist()))
    pauli_sets.append(([2, 4], [3, 3]))  # Z-only
    pauli_sets.append(([1], [2]))        # single Y
    for tg, pl in pauli_sets:
        cases.append(("pauli", "pauli", tg, [], pl, 0.0, None))
        cases.append(("pauli_rot", "pauli_rot", tg, [], pl, float(rng.uniform(0, 2 * np.pi)), None))

    names, kinds, tgs, cls, cvs, params, mats, ins, outs = [], [], [], [], [], [], [], [], []
    for (name, kind, tg, cl, cv, param, mat) in cases:
        psi = random_state(rng, n)
        out = psi.copy()
        if kind == "dense":
            if len(tg) == 1:
                ref.dense1(tg[0], mat, out)
            elif len(tg) == 2:
                ref.dense2(tg[0], tg[1], mat, out)
            else:
                ref.dense(tg, mat, out)
        elif kind == "cdense":
            if len(tg) == 1 and len(cl) == 1:
                m = np.ascontiguousarray(mat, dtype=np.complex128)
                ref.f("single_qubit_control_single_qubit_dense_matrix_gate")(
                    O._u(cl[0]), O._u(cv[0]), O._u(tg[0]), O._ptr(m), O._ptr(out), O._ull(dim))
            else:
                ref.controlled_dense(cl, cv, tg, mat, out)
        elif kind == "named":
            ref.named(name, tg[0], out)
        elif kind == "rot":
            ref.rotation(name, tg[0], param, out)
        elif kind == "two":
            ref.two(name, cl[0], tg[0], out)
        elif kind == "diag1":
            ref.diag1(tg[0], mat, out)
        elif kind == "phase1":
            ref.phase1(tg[0], complex(mat[0]), out)
        elif kind == "cdiag":
            if cl:
                ref.controlled_diag(cl, cv, tg, mat, out)
            else:
                tt = O._u32(tg)
                dd = O._c128(mat)
                ref.f("multi_qubit_diagonal_matrix_gate")(O._ptr(tt), O._u(tt.size), O._ptr(dd), O._ptr(out), O._ull(dim))
        elif kind == "pauli":
            ref.pauli(tg, cv, out)
        elif kind == "pauli_rot":
            ref.pauli_rotation(tg, cv, param, out)
        names.append(name)
        kinds.append(kind)
        tgs.append(np.array(tg + [-1] * (8 - len(tg)), dtype=np.int64))
        cls.append(np.array(cl + [-1] * (8 - len(cl)), dtype=np.int64))
        cvs.append(np.array(cv + [-1] * (8 - len(cv)), dtype=np.int64))
        params.append(param)
        mats.append(np.zeros(0, dtype=np.complex128) if mat is None else np.asarray(mat, dtype=np.complex128).ravel())
        ins.append(psi)
        outs.append(out)
    np.savez_compressed(OUT / "sv_gates.npz", n=n, name=np.array(names), kind=np.array(kinds),
                        targets=np.array(tgs), controls=np.array(cls), cvals=np.array(cvs),
                        param=np.array(params), matrix=np.concatenate(mats),
                        matrix_offset=np.cumsum([0] + [m.size for m in mats]), psi_in=np.array(ins),
                        psi_out=np.array(outs))

    # ------------------------------------------------------------------ reductions
    psi = random_state(rng, n)
    bra = random_state(rng, n)
    terms = random_pauli_terms(rng, n, 24)
    exp_vals = np.array([ref.expectation(t, p, psi) for (t, p, _c) in terms])
    trans = np.array([ref.transition(t, p, bra, psi) for (t, p, _c) in terms])
    marg_t = [[0, 3, 5], [1], [2, 6]]
    marg_v = [[1, 0, 1], [0], [1, 1]]
    margs = np.array([ref.marginal(t, v, psi) for t, v in zip(marg_t, marg_v)])
    tt = np.full((len(terms), n), -1, dtype=np.int64)
    pp = np.full((len(terms), n), -1, dtype=np.int64)
    for i, (t, p, _c) in enumerate(terms):
        tt[i, : len(t)] = t
        pp[i, : len(p)] = p
    np.savez_compressed(
        OUT / "sv_stats.npz", n=n, psi=psi, bra=bra, term_targets=tt, term_paulis=pp,
        expectation=exp_vals, transition=trans,
        norm=ref.norm_squared(psi), inner=ref.inner_product(bra, psi),
        M0=np.array([ref.M0(t, psi) for t in range(n)]), M1=np.array([ref.M1(t, psi) for t in range(n)]),
        entropy=ref.entropy(psi), marg_targets=np.array([t + [-1] * (3 - len(t)) for t in marg_t]),
        marg_values=np.array([v + [-1] * (3 - len(v)) for v in marg_v]), marginal=margs)

    # ------------------------------------------------------------------ density matrix
    nd = 4
    ddim = 1 << nd
    dm_cases = []
    for t in range(nd):
        dm_cases.append(("dm_dense1", [t], [], [], haar_unitary(rng, 1)))
    for tg in ([0, 1], [3, 1], [2, 0, 3]):
        dm_cases.append(("dm_dense", tg, [], [], haar_unitary(rng, len(tg))))
    dm_cases.append(("dm_cdense", [2], [0], [1], haar_unitary(rng, 1)))
    dm_cases.append(("dm_cdense", [1, 3], [2], [0], haar_unitary(rng, 2)))
    for nm in ("X", "Y", "Z", "H", "S", "T", "P0", "P1"):
        dm_cases.append((f"dm_{nm}", [1], [], [], None))
    for nm in ("CNOT", "CZ", "SWAP"):
        dm_cases.append((f"dm_{nm}", [3], [0], [], None))
    for nm in ("RX", "RY", "RZ"):
        dm_cases.append((f"dm_{nm}", [2], [], [], float(rng.uniform(0, 2 * np.pi))))
    dnames, dtg, dcl, dcv, dmat, dpar, din, dout = [], [], [], [], [], [], [], []
    for (name, tg, cl, cv, mat) in dm_cases:
        # mixed state: a random mixture of three pure states
        rho = np.zeros(ddim * ddim, dtype=np.complex128)
        for w in rng.dirichlet(np.ones(3)):
            rho += w * ref.dm_from_pure(random_state(rng, nd))
        out = rho.copy()
        param = 0.0
        if name == "dm_dense1":
            ref.dm_dense1(tg[0], mat, out, ddim)
        elif name == "dm_dense":
            ref.dm_dense(tg, mat, out, ddim)
        elif name == "dm_cdense":
            if len(tg) == 1:
                c, v, m = O._u32(cl), O._u32(cv), O._c128(mat)
                ref.f("dm_multi_qubit_control_single_qubit_dense_matrix_gate")(
                    O._ptr(c), O._ptr(v), O._u(c.size), O._u(tg[0]), O._ptr(m), O._ptr(out), O._ull(ddim))
            else:
                ref.dm_controlled_dense(cl, cv, tg, mat, out, ddim)
        elif name in ("dm_CNOT", "dm_CZ", "dm_SWAP"):
            ref.dm_two(name[3:], cl[0], tg[0], out, ddim)
        elif name in ("dm_RX", "dm_RY", "dm_RZ"):
            param = mat
            mat = None
            ref.dm_rotation(name[3:], tg[0], param, out, ddim)
        else:
            ref.dm_named(name[3:], tg[0], out, ddim)
        dnames.append(name)
        dtg.append(np.array(tg + [-1] * (4 - len(tg))))
        dcl.append(np.array(cl + [-1] * (4 - len(cl))))
        dcv.append(np.array(cv + [-1] * (4 - len(cv))))
        mm = np.zeros((8, 8), dtype=np.complex128)
        if mat is not None:
            a = np.asarray(mat, dtype=np.complex128)
            mm[: a.shape[0], : a.shape[1]] = a
        dmat.append(mm)
        dpar.append(param)
        din.append(rho)
        dout.append(out)
    rho = np.zeros(ddim * ddim, dtype=np.complex128)
    for w in rng.dirichlet(np.ones(3)):
        rho += w * ref.dm_from_pure(random_state(rng, nd))
    dterms = random_pauli_terms(rng, nd, 12)
    dtt = np.full((len(dterms), nd), -1, dtype=np.int64)
    dpp = np.full((len(dterms), nd), -1, dtype=np.int64)
    for i, (t, p, _c) in enumerate(dterms):
        dtt[i, : len(t)] = t
        dpp[i, : len(p)] = p
    np.savez_compressed(
        OUT / "dm.npz", n=nd, name=np.array(dnames), targets=np.array(dtg), controls=np.array(dcl),
        cvals=np.array(dcv), matrix=np.array(dmat), param=np.array(dpar), rho_in=np.array(din), rho_out=np.array(dout),
        rho=rho, term_targets=dtt, term_paulis=dpp,
        expectation=np.array([ref.dm_expectation(t, p, rho, ddim) for (t, p, _c) in dterms]),
        trace=ref.dm_norm(rho, ddim), M0=np.array([ref.dm_M0(t, rho, ddim) for t in range(nd)]),
        marginal=ref.dm_marginal([0, 2], [1, 0], rho, ddim))
    print("wrote", sorted(p.name for p in OUT.glob("*.npz")))


if __name__ == "__main__":
    main()
