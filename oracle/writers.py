"""CPU restatement of hilbert's RDM file writers (test infrastructure only).

Only tests use this module, to check the library's files; the product never
calls it.

What it restates, record by record and in writing order, from one instance's
primal vector of an OracleProblem (orbitals in Pitzer order):

* ``WriteOPDM`` (v2rdm_casscf/write_opdm.cc:40-119): D1a / D1b records
  {i, j, value}, active-active by irrep (:57-84), then core-core = 1 (:87-106).
* ``WriteTPDM`` (v2rdm_casscf/write_tpdm.cc:40-300): D2aa / D2bb / D2ab
  records {i, j, k, l, value} = D2(ij; kl) over the geminal pairs of every
  irrep (:60-103), core-core (:106-148), core-active from the 1-RDM
  (:151-256), and the ACTIVE / INACTIVE ORBITALS lists (:262-300).
* ``WriteActiveTPDM`` (write_tpdm.cc:521-595): the active-active D2 records
  of WriteTPDM alone.
* ``WriteTPDMSpinFree`` (write_tpdm.cc:312-519): spin-free D2 records with
  hilbert's same-spin sign rule (see ``spin_free_records``).
* ``WriteTPDM_IWL`` (v2rdm_casscf/write_tpdm_iwl.cc:45-200): the same
  elements with labels (i, k, j, l) and the aa / bb values scaled by 1/4.

Orbital labels are full-space: ``active[p]`` for active orbital p (Pitzer
position) and the ``core`` list in hilbert's order.

The binary containers are read back by ``read_keyed`` / ``read_iwl``:
a keyed entry is an 80-byte key, an int64 byte count and the bytes; the IWL
entry "IWL Buffers" holds psi4 libiwl buffers (2980 values each: int lastbuf,
int inbuf, int16 labels[4 * 2980], double values[2980]).
"""
from __future__ import annotations

import numpy as np

OPDM = np.dtype([("i", "<i4"), ("j", "<i4"), ("value", "<f8")])
TPDM = np.dtype([("i", "<i4"), ("j", "<i4"), ("k", "<i4"), ("l", "<i4"), ("value", "<f8")])
IWL_INTS_PER_BUF = 2980
IWL_CUTOFF = 1.0e-14


class _Tables:
    def __init__(self, P):
        t = P.tables()
        off = P.offsets()
        self.H = P.nirrep
        self.gab, self.gaa = t["gems_ab"], t["gems_aa"]
        starts = np.concatenate([[0], np.cumsum(self.gab)])
        self.bas_ab = [t["bas_ab"][starts[h]:starts[h + 1]] for h in range(self.H)]
        self.ibas_aa = t["ibas_aa"]
        self.off = off
        self.amopi = np.bincount(np.asarray(P.irreps), minlength=self.H)
        self.poff = np.concatenate([[0], np.cumsum(self.amopi)])[:self.H]


def _active_pairs(T, x, active):
    """(i, j, k, l, ab, same, aa, bb) over the geminal pairs
    (write_tpdm.cc:60-103, write_tpdm_iwl.cc:56-95)."""
    for h in range(T.H):
        g, ga = int(T.gab[h]), int(T.gaa[h])
        for ij in range(g):
            i, j = T.bas_ab[h][ij]
            for kl in range(g):
                k, l = T.bas_ab[h][kl]
                vab = x[T.off["d2ab"][h] + ij * g + kl]
                same = i != j and k != l
                vaa = vbb = 0.0
                if same:
                    ija, kla = T.ibas_aa[h, i, j], T.ibas_aa[h, k, l]
                    s = (1 if i < j else -1) * (1 if k < l else -1)
                    vaa = s * x[T.off["d2aa"][h] + ija * ga + kla]
                    vbb = s * x[T.off["d2bb"][h] + ija * ga + kla]
                yield active[i], active[j], active[k], active[l], vab, same, vaa, vbb


def _core_active(T, x, active, core):
    """(i, j, l, D1a(j, l), D1b(j, l)) for core i and active j, l of one irrep
    (write_tpdm.cc:151-256)."""
    for ci in core:
        for h in range(T.H):
            m = int(T.amopi[h])
            for j in range(m):
                for l in range(m):
                    va = x[T.off["d1a"][h] + j * m + l]
                    vb = x[T.off["d1b"][h] + j * m + l]
                    yield ci, active[T.poff[h] + j], active[T.poff[h] + l], va, vb


def opdm_records(P, x, active, core, beta):
    """D1a (beta False) or D1b records (write_opdm.cc:57-106)."""
    T = _Tables(P)
    key = "d1b" if beta else "d1a"
    out = []
    for h in range(T.H):
        m = int(T.amopi[h])
        for i in range(m):
            for j in range(m):
                out.append((active[T.poff[h] + i], active[T.poff[h] + j], x[T.off[key][h] + i * m + j]))
    out += [(c, c, 1.0) for c in core]
    return np.array(out, dtype=OPDM)


def tpdm_records(P, x, active, core, part, active_only=False):
    """D2aa (part 0), D2bb (1) or D2ab (2) records: WriteTPDM
    (write_tpdm.cc:60-256) or, active_only, WriteActiveTPDM (:536-584)."""
    T = _Tables(P)
    out = []
    for i, j, k, l, vab, same, vaa, vbb in _active_pairs(T, x, active):
        if part == 2:
            out.append((i, j, k, l, vab))
        elif same:
            out.append((i, j, k, l, vbb if part == 1 else vaa))
    if active_only:
        return np.array(out, dtype=TPDM)
    for ci in core:
        for cj in core:
            if part == 2:
                out.append((ci, cj, ci, cj, 1.0))
            elif ci != cj:
                out += [(ci, cj, ci, cj, 1.0), (ci, cj, cj, ci, -1.0)]
    for i, j, l, va, vb in _core_active(T, x, active, core):
        if part == 2:
            out += [(i, j, i, l, vb), (j, i, l, i, va)]
        else:
            v = vb if part == 1 else va
            out += [(i, j, i, l, v), (i, j, l, i, -v), (j, i, l, i, v), (j, i, i, l, -v)]
    return np.array(out, dtype=TPDM)


def spin_free_records(P, x, active, core, hilbert_sign=True):
    """Spin-free D2 records of WriteTPDMSpinFree (write_tpdm.cc:326-486):
    D2ab(ij;kl) + D2ab(ji;lk) + sg (D2aa + D2bb)(ij;kl) over the stored
    same-spin pair.  hilbert sets sg = -1 as soon as i > j or k > l
    (write_tpdm.cc:356-358), so also when both hold; hilbert_sign=False takes
    the product of the two pair signs instead (the spin-free 2-RDM proper)."""
    T = _Tables(P)
    out = []
    for h in range(T.H):
        g, ga = int(T.gab[h]), int(T.gaa[h])
        pos = {(int(a), int(b)): t for t, (a, b) in enumerate(T.bas_ab[h])}
        for ij in range(g):
            i, j = T.bas_ab[h][ij]
            ji = pos[(int(j), int(i))]
            for kl in range(g):
                k, l = T.bas_ab[h][kl]
                lk = pos[(int(l), int(k))]
                v = x[T.off["d2ab"][h] + ij * g + kl] + x[T.off["d2ab"][h] + ji * g + lk]
                if i != j and k != l:
                    ija, kla = T.ibas_aa[h, i, j], T.ibas_aa[h, k, l]
                    if hilbert_sign:
                        sg = -1.0 if (i > j or k > l) else 1.0
                    else:
                        sg = (1.0 if i < j else -1.0) * (1.0 if k < l else -1.0)
                    v += sg * x[T.off["d2aa"][h] + ija * ga + kla]
                    v += sg * x[T.off["d2bb"][h] + ija * ga + kla]
                out.append((active[i], active[j], active[k], active[l], v))
    for ci in core:
        for cj in core:
            out.append((ci, cj, ci, cj, 4.0 if ci != cj else 2.0))
            if ci != cj:
                out.append((ci, cj, cj, ci, -2.0))
    for i, j, l, va, vb in _core_active(T, x, active, core):
        d, e = va + vb + vb + va, -va - vb
        out += [(i, j, i, l, d), (i, j, l, i, e), (j, i, l, i, d), (j, i, i, l, e)]
    return np.array(out, dtype=TPDM)


def iwl_records(P, x, active, core, part):
    """IWL values of part 0 aa, 1 bb, 2 ab as written, before the cutoff
    (write_tpdm_iwl.cc:56-183)."""
    T = _Tables(P)
    out = []
    for i, j, k, l, vab, same, vaa, vbb in _active_pairs(T, x, active):
        if part == 2:
            out.append((i, k, j, l, vab))
        elif same:
            out.append((i, k, j, l, 0.25 * (vbb if part == 1 else vaa)))
    for ci in core:
        for cj in core:
            if part == 2:
                out.append((ci, ci, cj, cj, 1.0))
            elif ci != cj:
                out += [(ci, ci, cj, cj, 0.25), (ci, cj, cj, ci, -0.25)]
    for i, j, l, va, vb in _core_active(T, x, active, core):
        if part == 2:
            out += [(i, i, j, l, vb), (j, l, i, i, va)]
        else:
            v = 0.25 * (vb if part == 1 else va)
            out += [(i, i, j, l, v), (i, l, j, i, -v), (j, l, i, i, v), (j, i, i, l, -v)]
    return np.array(out, dtype=TPDM)


def read_keyed(path):
    """{key: bytes} of a file of keyed entries, in file order."""
    raw = open(path, "rb").read()
    out, pos = {}, 0
    while pos < len(raw):
        key = raw[pos:pos + 80].split(b"\0", 1)[0].decode()
        n = int(np.frombuffer(raw, dtype="<i8", count=1, offset=pos + 80)[0])
        out[key] = raw[pos + 88:pos + 88 + n]
        pos += 88 + n
    return out


def read_iwl(path):
    """(records, buffers) of an IWL file: the values in writing order and the
    (lastbuf, inbuf) header of every buffer."""
    blob = read_keyed(path)["IWL Buffers"]
    n = IWL_INTS_PER_BUF
    size = 8 + 8 * n + 8 * n
    recs, heads = [], []
    for b in range(len(blob) // size):
        q = blob[b * size:(b + 1) * size]
        last, inbuf = np.frombuffer(q, dtype="<i4", count=2)
        lab = np.frombuffer(q, dtype="<i2", count=4 * n, offset=8).reshape(n, 4)
        val = np.frombuffer(q, dtype="<f8", count=n, offset=8 + 8 * n)
        heads.append((int(last), int(inbuf)))
        for t in range(inbuf):
            recs.append((*lab[t].tolist(), val[t]))
    if len(blob) % size:
        raise ValueError("IWL entry is not a whole number of buffers")
    return np.array(recs, dtype=TPDM), heads
