"""The reference's determinant bookkeeping for file04 CI vectors (test infrastructure only).

* ``jdifgen`` restates Jdifgen (koopro4QD.F90:3763-3886): the spin-orbitals in which two determinants,
  each a list of holes and particles relative to the closed-shell reference, differ, and the parity ZSIG
  of the crossings needed to align their lists.  For determinants II and JJ that differ in one
  spin-orbital it gives <II| a+_{NV2} a_{NV1} |JJ> = (-1)^ZSIG.
* ``one_rdm`` builds the spin-free 1-RDM of a file04 vector from those matrix elements alone, without
  ever ordering a determinant: an independent check of how nv2_ci_from_file04 phases each determinant.
* ``lists_from_ci`` writes a CI vector in synth's string layout as file04 excitation lists, with the
  holes and particles of each determinant in a shuffled order and the coefficient carrying the phase
  that order implies (pairs a+_part a_trou applied in listed order), for test inputs.
"""
from __future__ import annotations

from itertools import combinations

import numpy as np


def jdifgen(ne, nd, trou, part, II, JJ, nma=4):
    """Restates Jdifgen (0-based II, JJ; spin-orbitals 1-based as in file04).  Returns
    (ok, nbdif, nv1, nv2, zsig); ok is False when the determinants differ in more than nma spin-orbitals."""
    KR, KL = II, JJ
    zshift = False
    if ne[KR] < ne[KL]:
        KR, KL = KL, KR
        zshift = True
    NE1, NE2 = int(ne[KR]), int(ne[KL])
    if NE1 == 0:
        raise ValueError("identical determinants")
    nbdif = NE1 - NE2
    if nbdif > nma:
        return False, nbdif, [], [], False
    ita = [int(trou[nd[KR] + j]) for j in range(NE1)] + [int(part[nd[KR] + j]) for j in range(NE1)]
    zsig = False
    nv1, nv2, natu1, natu2 = [0] * (nma + 1), [0] * (nma + 1), [0] * (nma + 1), [0] * (nma + 1)
    if NE2 > 0:
        itb = [int(trou[nd[KL] + j]) for j in range(NE2)] + [int(part[nd[KL] + j]) for j in range(NE2)]
        for K in (1, 2):
            K1, K2 = (K - 1) * NE1, (K - 1) * NE2
            for I in range(1, NE2 + 1):
                nit1 = itb[K2 + I - 1]
                found = False
                for J in range(1 + K1, NE1 + K1 + 1):
                    njt2 = ita[J - 1]
                    if nit1 != njt2:
                        continue
                    found = True
                    if I != J - K1:
                        ita[J - 1] = ita[I + K1 - 1]
                        ita[I + K1 - 1] = njt2
                        zsig = not zsig
                    break
                if not found:
                    nbdif += 1
                    if nbdif > nma:
                        return False, nbdif, [], [], zsig
                    nv1[nbdif] = nit1
                    natu1[nbdif] = K
                    natu2[nbdif] = I
    if nbdif <= 0:
        raise ValueError("identical determinants")
    nar = NE1 - NE2
    for I in range(1, nar + 1):
        net = NE1 + 1 - I
        nv1[I] = ita[net - 1]
        nv2[I] = ita[net + NE1 - 1]
    for I in range(nar + 1, nbdif + 1):
        K, NI, NU = natu1[I], natu2[I], nv1[I]
        nv2[I] = ita[NI + NE1 * (K - 1) - 1]
        if K == 2:
            continue
        nv1[I] = nv2[I]
        nv2[I] = NU
        zsig = not zsig
    a, b = nv1[1:nbdif + 1], nv2[1:nbdif + 1]
    if zshift:
        a, b = b, a
    return True, nbdif, a, b, zsig


def one_rdm(f04, ncore, nact, state=0):
    """Spin-free gamma(p,q) = <E_pq> over the active orbitals from the file04 lists and Jdifgen."""
    norb, noca, ncf = f04["norb"], f04["noca"], f04["ncf"]
    ne, nd, trou, part = f04["ne"], f04["nd"], f04["trou"], f04["part"]
    c = f04["c"][:, state]
    g = np.zeros((nact, nact))
    spat = lambda so: (so - 1) % norb
    for m in range(ncf):
        occ = np.zeros(2 * norb + 1)
        occ[:noca] = 1
        occ[norb:norb + noca] = 1
        for k in range(ne[m]):
            occ[trou[nd[m] + k] - 1] -= 1
            if part[nd[m] + k] <= 2 * norb:
                occ[part[nd[m] + k] - 1] += 1
        for a in range(nact):
            g[a, a] += c[m] ** 2 * (occ[ncore + a] + occ[norb + ncore + a])
        for n in range(ncf):
            if n == m:
                continue
            ok, nbdif, nv1, nv2, zsig = jdifgen(ne, nd, trou, part, m, n, 1)
            if not ok or nbdif != 1:
                continue
            x, y = nv2[0], nv1[0]                    # <m| a+_x a_y |n> = (-1)^zsig
            if (x - 1) // norb != (y - 1) // norb:
                continue
            p, q = spat(x) - ncore, spat(y) - ncore
            g[p, q] += c[m] * c[n] * (-1.0 if zsig else 1.0)
    return g


def lists_from_ci(C, nact, ncore, norb, nalpha, nbeta, seed=0, grouped=False):
    """file04 lists for C[root, I_alpha, I_beta] (synth's strings; alpha = spin-orbitals 1..norb).
    The reference determinant closes ncore + min(nalpha, nbeta) orbitals; each determinant's holes and
    particles are shuffled, and its coefficient carries the phase of that order.  ``grouped`` phases the
    determinants as a+_{p1} .. a+_{pk} a_{t1} .. a_{tk} |ref> instead (a planted convention error)."""
    rng = np.random.default_rng(seed)
    nocc = min(nalpha, nbeta)
    noca = ncore + nocc
    ref = set(range(1, noca + 1)) | set(range(norb + 1, norb + noca + 1))
    ne, nd, trou, part, coefs = [], [], [], [], []
    sa = list(combinations(range(nact), nalpha))
    sb = list(combinations(range(nact), nbeta))
    core = set(range(1, ncore + 1)) | set(range(norb + 1, norb + ncore + 1))
    for ia, a in enumerate(sa):
        for ib, b in enumerate(sb):
            occ = set(core) | {ncore + 1 + x for x in a} | {norb + ncore + 1 + x for x in b}
            holes = sorted(ref - occ)
            parts = sorted(occ - ref)
            if len(holes) != len(parts):
                raise ValueError("open shells beyond the reference need the ion convention")
            rng.shuffle(holes)
            rng.shuffle(parts)
            # phase of prod_k a+_{p_k} a_{t_k} |ref> relative to the ascending order of occ
            bits = set(ref)
            sign = 1
            ops = ([("a", t) for t in reversed(holes)] + [("c", p) for p in reversed(parts)]) if grouped \
                else [op for t, p in zip(holes, parts) for op in (("a", t), ("c", p))]
            for kind, o in ops:             # rightmost operator of the product first
                if sum(1 for x in bits if x < o) & 1:
                    sign = -sign
                if kind == "a":
                    bits.remove(o)
                else:
                    bits.add(o)
            nd.append(len(trou))
            ne.append(len(holes))
            trou += holes
            part += parts
            coefs.append(sign * C[:, ia, ib])
    return {"norb": norb, "noca": noca, "nocb": noca, "metat": C.shape[0], "ncf": len(ne),
            "ne": np.array(ne, dtype=np.int16), "nd": np.array(nd, dtype=np.int64),
            "trou": np.array(trou, dtype=np.int16), "part": np.array(part, dtype=np.int16),
            "c": np.array(coefs), "e": np.zeros(C.shape[0])}
