"""Brute-force second quantisation for small active spaces (test helper).

Spin orbitals: alpha p -> mode p, beta p -> mode n + p.  Jordan-Wigner
sign convention: a_k |...> picks up (-1)^(number of occupied modes < k).
Used to build exact wavefunctions and their RDMs, independently of the
oracle's maps, so that A x(Psi) = b and c.x(Psi) = <Psi|H|Psi> pin them.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class Fock:
    def __init__(self, n):
        self.n = n
        self.M = 2 * n
        dim = 1 << self.M
        self.dim = dim
        idx = np.arange(dim)
        self.occ = ((idx[:, None] >> np.arange(self.M)[None, :]) & 1).astype(np.int8)
        self._ann = []
        for k in range(self.M):
            has = self.occ[:, k] == 1
            src = idx[has]
            dst = src ^ (1 << k)
            sign = (-1.0) ** self.occ[src, :k].sum(axis=1)
            self._ann.append(sp.csr_matrix((sign, (dst, src)), shape=(dim, dim)))

    def a(self, p, s):
        return self._ann[p + s * self.n]

    def ad(self, p, s):
        return self._ann[p + s * self.n].T.tocsr()

    def number(self, s):
        return self.occ[:, s * self.n:(s + 1) * self.n].sum(axis=1)

    def irrep(self, irreps):
        g = np.zeros(self.dim, dtype=np.int64)
        irr = np.asarray(list(irreps) + list(irreps))
        for k in range(self.M):
            g ^= np.where(self.occ[:, k] == 1, irr[k], 0)
        return g

    def sector(self, na, nb, irreps=None, gamma=None):
        sel = (self.number(0) == na) & (self.number(1) == nb)
        if irreps is not None and gamma is not None:
            sel &= self.irrep(irreps) == gamma
        return np.nonzero(sel)[0]

    def hamiltonian(self, h, g):
        """H = sum h_pq a+_ps a_qs + 1/2 sum (pq|rs) a+_ps a+_rt a_st a_qs."""
        n = self.n
        H = sp.csr_matrix((self.dim, self.dim))
        for s in range(2):
            for p in range(n):
                for q in range(n):
                    if h[p, q] != 0.0:
                        H = H + h[p, q] * (self.ad(p, s) @ self.a(q, s))
        for s, t in itertools.product(range(2), range(2)):
            for p, q, r, u in itertools.product(range(n), repeat=4):
                v = g[p, q, r, u]
                if v == 0.0:
                    continue
                op = self.ad(p, s) @ self.ad(r, t) @ self.a(u, t) @ self.a(q, s)
                H = H + 0.5 * v * op
        return H.tocsr()

    def s_squared(self):
        n = self.n
        Sp = sum(self.ad(p, 0) @ self.a(p, 1) for p in range(n))
        Sm = Sp.T
        Sz = 0.5 * sp.diags((self.number(0) - self.number(1)).astype(float))
        return (Sm @ Sp + Sz @ Sz + Sz).tocsr()


def spin_adapted_random_state(F, na, nb, rng, irreps=None, gamma=0):
    """Random real state in the (na, nb[, gamma]) sector with S = Ms = (na-nb)/2."""
    sec = F.sector(na, nb, irreps, gamma if irreps is not None else None)
    S2 = F.s_squared()[sec][:, sec].toarray()
    w, v = np.linalg.eigh(S2)
    ms = 0.5 * (na - nb)
    keep = np.abs(w - ms * (ms + 1)) < 1e-8
    assert keep.any(), "no S = Ms states in sector"
    coef = v[:, keep] @ rng.standard_normal(keep.sum())
    psi = np.zeros(F.dim)
    psi[sec] = coef / np.linalg.norm(coef)
    return psi


def ground_state(F, H, na, nb, irreps=None, gamma=None):
    sec = F.sector(na, nb, irreps, gamma)
    Hs = H[sec][:, sec]
    if len(sec) <= 2000:
        w, v = np.linalg.eigh(Hs.toarray())
        e, c = w[0], v[:, 0]
    else:
        w, v = spla.eigsh(Hs, k=1, which="SA")
        e, c = w[0], v[:, 0]
    psi = np.zeros(F.dim)
    psi[sec] = c
    return e, psi


def lowest_spin_state(F, H, na, nb, irreps=None, gamma=None):
    """Lowest eigenpair of H within the S = Ms = (na-nb)/2 subspace of a sector."""
    sec = F.sector(na, nb, irreps, gamma)
    S2 = F.s_squared()[sec][:, sec].toarray()
    w, v = np.linalg.eigh(S2)
    ms = 0.5 * (na - nb)
    B = v[:, np.abs(w - ms * (ms + 1)) < 1e-8]
    Hs = B.T @ (H[sec][:, sec] @ B)
    e, c = np.linalg.eigh(Hs)
    psi = np.zeros(F.dim)
    psi[sec] = B @ c[:, 0]
    return e[0], psi


class RDMs:
    """All 1- and 2-body expectation values needed for the DQG primal."""

    def __init__(self, F, psi):
        self.F = F
        n = F.n
        M = F.M
        ann = [F._ann[k] @ psi for k in range(M)]
        cre = [F._ann[k].T @ psi for k in range(M)]
        # D1[p,q] = <a+_p a_q> = (a_p psi).(a_q psi)
        self.D1 = np.array([[ann[p] @ ann[q] for q in range(M)] for p in range(M)])
        # two-annihilated vectors v_(r,s) = a_r a_s psi
        V = np.array([F._ann[r] @ ann[s] for r in range(M) for s in range(M)])
        G = V @ V.T  # G[(q,p),(r,s)] = <a+_p a+_q a_r a_s>
        self.D2 = G.reshape(M, M, M, M).transpose(1, 0, 2, 3)  # D2[p,q,r,s] = <a+_p a+_q a_r a_s>
        C2 = np.array([F._ann[r].T @ cre[s] for r in range(M) for s in range(M)])
        Q = C2 @ C2.T  # Q[(q,p),(r,s)] = <a_q a_p a+_r a+_s> -> <a_p a_q a+_r a+_s> via transpose
        self.Q2 = Q.reshape(M, M, M, M).transpose(1, 0, 2, 3)  # <a_p a_q a+_r a+_s>
        W = np.array([F._ann[r].T @ ann[s] for r in range(M) for s in range(M)])  # a+_r a_s psi
        Gm = W @ W.T  # <(a+_j a_i)^+ ...>: Gm[(j,i),(l,k)] = <a+_i a_j a+_l a_k>
        self.G = Gm.reshape(M, M, M, M).transpose(1, 0, 2, 3)  # G[i,j,l,k] = <a+_i a_j a+_l a_k>
        self.Q1 = np.eye(M) - self.D1.T  # <a_p a+_q> = delta - <a+_q a_p>
        self.n = n

    def m(self, p, s):
        return p + s * self.n


def pack_primal(P, R):
    """Pack exact RDMs into the oracle's primal layout (P: OracleProblem)."""
    t = P.tables()
    off = P.offsets()
    n, H = P.nmo, P.nirrep
    irr = np.asarray(P.irreps)
    amopi = np.bincount(irr, minlength=H)
    poff = np.concatenate([[0], np.cumsum(amopi)[:-1]])
    gab, gaa = t["gems_ab"], t["gems_aa"]
    bab_start = np.concatenate([[0], np.cumsum(gab)[:-1]])
    baa_start = np.concatenate([[0], np.cumsum(gaa)[:-1]])
    x = np.zeros(P.n_primal)
    m = R.m
    for h in range(H):
        mh = amopi[h]
        for i in range(mh):
            for j in range(mh):
                ii, jj = i + poff[h], j + poff[h]
                x[off["d1a"][h] + i * mh + j] = R.D1[m(ii, 0), m(jj, 0)]
                x[off["d1b"][h] + i * mh + j] = R.D1[m(ii, 1), m(jj, 1)]
                x[off["q1a"][h] + i * mh + j] = R.Q1[m(ii, 0), m(jj, 0)]
                x[off["q1b"][h] + i * mh + j] = R.Q1[m(ii, 1), m(jj, 1)]
        g = gab[h]
        pairs = t["bas_ab"][bab_start[h]:bab_start[h] + g]
        for a, (i, j) in enumerate(pairs):
            for b_, (k, l) in enumerate(pairs):
                e = a * g + b_
                x[off["d2ab"][h] + e] = R.D2[m(i, 0), m(j, 1), m(l, 1), m(k, 0)]
                if P.flags[0]:
                    x[off["q2ab"][h] + e] = R.Q2[m(i, 0), m(j, 1), m(l, 1), m(k, 0)]
                if P.flags[1]:
                    x[off["g2ab"][h] + e] = R.G[m(i, 0), m(j, 1), m(l, 1), m(k, 0)]
                    x[off["g2ba"][h] + e] = R.G[m(i, 1), m(j, 0), m(l, 0), m(k, 1)]
                    G2 = 2 * g
                    base = off["g2aa"][h]
                    x[base + a * G2 + b_] = R.G[m(i, 0), m(j, 0), m(l, 0), m(k, 0)]
                    x[base + (a + g) * G2 + b_ + g] = R.G[m(i, 1), m(j, 1), m(l, 1), m(k, 1)]
                    x[base + a * G2 + b_ + g] = R.G[m(i, 0), m(j, 0), m(l, 1), m(k, 1)]
                    x[base + (a + g) * G2 + b_] = R.G[m(i, 1), m(j, 1), m(l, 0), m(k, 0)]
        ga = gaa[h]
        apairs = t["bas_aa"][baa_start[h]:baa_start[h] + ga]
        for a, (i, j) in enumerate(apairs):
            for b_, (k, l) in enumerate(apairs):
                e = a * ga + b_
                x[off["d2aa"][h] + e] = R.D2[m(i, 0), m(j, 0), m(l, 0), m(k, 0)]
                x[off["d2bb"][h] + e] = R.D2[m(i, 1), m(j, 1), m(l, 1), m(k, 1)]
                if P.flags[0]:
                    x[off["q2aa"][h] + e] = R.Q2[m(i, 0), m(j, 0), m(l, 0), m(k, 0)]
                    x[off["q2bb"][h] + e] = R.Q2[m(i, 1), m(j, 1), m(l, 1), m(k, 1)]
        if P.flags[2]:
            # spin-adapted D200 blocks from D2ab (the spin rows define them)
            ibab = t["ibas_ab"][h]
            d2 = lambda a_, b_: x[off["d2ab"][h] + a_ * g + b_]
            singlet = P.nalpha == P.nbeta
            G2 = g if singlet else 2 * g
            for a, (i, j) in enumerate(pairs):
                ji = ibab[j, i]
                dij = np.sqrt(2.0) if i == j else 1.0
                for b_, (k, l) in enumerate(pairs):
                    lk = ibab[l, k]
                    dkl = np.sqrt(2.0) if k == l else 1.0
                    s4 = d2(a, b_) + d2(ji, b_) + d2(a, lk) + d2(ji, lk)
                    x[off["d200"][h] + a * G2 + b_] = 0.5 / (dij * dkl) * s4
                    if not singlet:
                        x[off["d200"][h] + a * G2 + b_ + g] = 0.5 / dij * (d2(a, b_) - d2(a, lk) + d2(ji, b_) - d2(ji, lk))
                        x[off["d200"][h] + (a + g) * G2 + b_] = 0.5 / dkl * (d2(a, b_) + d2(a, lk) - d2(ji, b_) - d2(ji, lk))
                        x[off["d200"][h] + (a + g) * G2 + b_ + g] = 0.5 * (d2(a, b_) - d2(ji, b_) - d2(a, lk) + d2(ji, lk))
    return x


def pack_t1(P, F, psi, x):
    """Fill the T1 blocks of x (oracle layout, P built with t1=True) with the
    exact T1(ijk,lmn) = <a+_i a+_j a+_k a_n a_m a_l> + <a_i a_j a_k a+_n a+_m a+_l>
    of psi; spins aaa, bbb, aab (i, j alpha, k beta), bba.  Triplets of irrep
    h in loop order i, j, k (aab: i < j; aaa: i < j < k), Pitzer orbitals."""
    import itertools as it
    n, M = F.n, F.M
    irr = np.asarray(P.irreps)
    ann = [F._ann[k] for k in range(M)]
    cre = [F._ann[k].T.tocsr() for k in range(M)]
    A3, C3 = {}, {}
    for a, b, c in it.product(range(M), repeat=3):
        A3[a, b, c] = ann[a] @ (ann[b] @ (ann[c] @ psi))
        C3[a, b, c] = cre[a] @ (cre[b] @ (cre[c] @ psi))
    off = P.t1_offsets()
    for h in range(P.nirrep):
        trips = [(i, j, k) for i in range(n) for j in range(n) for k in range(n) if irr[i] ^ irr[j] ^ irr[k] == h]
        aab = [t for t in trips if t[0] < t[1]]
        aaa = [t for t in aab if t[1] < t[2]]
        for name, tl, sp in (("t1aaa", aaa, (0, 0, 0)), ("t1bbb", aaa, (1, 1, 1)),
                             ("t1aab", aab, (0, 0, 1)), ("t1bba", aab, (1, 1, 0))):
            t = len(tl)
            for a, (i, j, k) in enumerate(tl):
                I, J, K = i + sp[0] * n, j + sp[1] * n, k + sp[2] * n
                for b, (l, m, q) in enumerate(tl):
                    L_, M_, N_ = l + sp[0] * n, m + sp[1] * n, q + sp[2] * n
                    x[off[name][h] + a * t + b] = A3[K, J, I] @ A3[N_, M_, L_] + C3[K, J, I] @ C3[N_, M_, L_]
    return x


def pack_t2(P, F, psi, x):
    """Fill the T2 blocks of x (oracle layout, P built with t2=True) with the
    exact T2(ijk,lmn) = <B_ijk B+_lmn> + <B+_lmn B_ijk>, B_ijk = a+_i a+_j a_k,
    of psi.  Spins: T2aab (a, a, b), T2bba (b, b, a); T2aaa is one block over
    the aab triplets with spins (a, a, a) followed by all triplets with
    spins (a, b, b); T2bbb likewise with (b, b, b) and (b, a, a)."""
    import itertools as it
    n, M = F.n, F.M
    irr = np.asarray(P.irreps)
    ann = [F._ann[k] for k in range(M)]
    cre = [F._ann[k].T.tocsr() for k in range(M)]
    Bv, Bd = {}, {}
    for a, b, c in it.product(range(M), repeat=3):
        Bv[a, b, c] = cre[a] @ (cre[b] @ (ann[c] @ psi))
        Bd[a, b, c] = cre[c] @ (ann[b] @ (ann[a] @ psi))
    off = P.t1_offsets()

    def val(p, q):
        return Bd[p] @ Bd[q] + Bv[q] @ Bv[p]

    for h in range(P.nirrep):
        trips = [(i, j, k) for i in range(n) for j in range(n) for k in range(n) if irr[i] ^ irr[j] ^ irr[k] == h]
        aab = [t for t in trips if t[0] < t[1]]
        for name, parts in (("t2aab", [(aab, (0, 0, 1))]), ("t2bba", [(aab, (1, 1, 0))]),
                            ("t2aaa", [(aab, (0, 0, 0)), (trips, (0, 1, 1))]),
                            ("t2bbb", [(aab, (1, 1, 1)), (trips, (1, 0, 0))])):
            rows = [tuple(o + s * n for o, s in zip(t, sp)) for tl, sp in parts for t in tl]
            d = len(rows)
            for a, p in enumerate(rows):
                for b, q in enumerate(rows):
                    x[off[name][h] + a * d + b] = val(p, q)
    return x


def pack_d3(P, F, psi, x, order="nml"):
    """Fill the D3 blocks of x (oracle layout, P built with d3=True) with
    the exact 3-RDM D3(ijk,lmn) = <a+_i a+_j a+_k a_n a_m a_l> of psi
    (order="lmn": <a+_i a+_j a+_k a_l a_m a_n>); spins aaa, bbb, aab
    (i, j alpha, k beta), bba; triplets as in pack_t1."""
    import itertools as it
    n, M = F.n, F.M
    irr = np.asarray(P.irreps)
    ann = [F._ann[k] for k in range(M)]
    A3 = {}
    for a, b, c in it.product(range(M), repeat=3):
        A3[a, b, c] = ann[a] @ (ann[b] @ (ann[c] @ psi))
    off = P.t1_offsets()
    for h in range(P.nirrep):
        trips = [(i, j, k) for i in range(n) for j in range(n) for k in range(n) if irr[i] ^ irr[j] ^ irr[k] == h]
        aab = [t for t in trips if t[0] < t[1]]
        aaa = [t for t in aab if t[1] < t[2]]
        for name, tl, sp in (("d3aaa", aaa, (0, 0, 0)), ("d3bbb", aaa, (1, 1, 1)),
                             ("d3aab", aab, (0, 0, 1)), ("d3bba", aab, (1, 1, 0))):
            t = len(tl)
            for a, (i, j, k) in enumerate(tl):
                I, J, K = i + sp[0] * n, j + sp[1] * n, k + sp[2] * n
                for b, (l, m, q) in enumerate(tl):
                    L_, M_, N_ = l + sp[0] * n, m + sp[1] * n, q + sp[2] * n
                    right = A3[N_, M_, L_] if order == "nml" else A3[L_, M_, N_]
                    # <a+_i a+_j a+_k X> = (a_k a_j a_i psi) . (X psi)
                    x[off[name][h] + a * t + b] = A3[K, J, I] @ right
    return x
