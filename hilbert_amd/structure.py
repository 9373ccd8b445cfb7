"""Host-only view of one problem structure: geminal tables, the constraint
matrix A with b (CSR), the cost-vector gather map and gather-table statistics.
No GPU is touched (v2h_structure_* in include/v2rdm_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class Structure:
    def __init__(self, nmo, nalpha, nbeta, irreps=None, nirrep=None, multiplicity=0, q2=True, g2=True,
                 spin=True, sz=True, t1=False, t2=False, d3=False):
        d, keep = N.make_desc(nmo, nalpha, nbeta, irreps, nirrep, multiplicity, q2, g2, spin, sz, t1, t2, d3)
        self.nirrep = d.nirrep
        self.nmo = nmo
        h = C.c_void_p()
        N.check(N.lib().v2h_structure_create(C.byref(d), C.byref(h)))
        self._h = h
        del keep
        np_, nd, nnz, nb = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        N.check(N.lib().v2h_structure_sizes(self._h, C.byref(np_), C.byref(nd), C.byref(nnz), C.byref(nb)))
        self.n_primal, self.n_dual, self.nnz, self.n_blocks = np_.value, nd.value, nnz.value, nb.value

    def __del__(self):
        if getattr(self, "_h", None):
            N.lib().v2h_structure_destroy(self._h)
            self._h = None

    def block_dims(self):
        d = np.zeros(self.n_blocks, dtype=np.int32)
        N.check(N.lib().v2h_structure_block_dims(self._h, d))
        return d

    def tables(self):
        n = self.nmo
        gab = np.zeros(self.nirrep, dtype=np.int32)
        gaa = np.zeros(self.nirrep, dtype=np.int32)
        # every ordered pair lies in exactly one irrep: sum gems_ab = n^2,
        # sum gems_aa = n (n - 1) / 2
        bab = np.zeros(2 * n * n, dtype=np.int32)
        baa = np.zeros(max(n * (n - 1), 1), dtype=np.int32)
        N.check(N.lib().v2h_structure_tables(self._h, gab, gaa, bab, baa))
        assert gab.sum() == n * n and gaa.sum() == n * (n - 1) // 2
        return {"gems_ab": gab, "gems_aa": gaa, "bas_ab": bab.reshape(-1, 2),
                "bas_aa": baa[:n * (n - 1)].reshape(-1, 2)}

    def constraints(self):
        """(row_ptr, col, val, b) of A."""
        ptr = np.zeros(self.n_dual + 1, dtype=np.int64)
        col = np.zeros(max(self.nnz, 1), dtype=np.int32)
        val = np.zeros(max(self.nnz, 1))
        b = np.zeros(self.n_dual)
        N.check(N.lib().v2h_structure_constraints(self._h, ptr, col, val, b))
        return ptr, col[:self.nnz], val[:self.nnz], b

    def matrix(self):
        import scipy.sparse as sp
        ptr, col, val, b = self.constraints()
        return sp.csr_matrix((val, col, ptr), shape=(self.n_dual, self.n_primal)), b

    def cost_map(self):
        i1 = np.zeros(self.n_primal, dtype=np.int32)
        i2 = np.zeros(self.n_primal, dtype=np.int32)
        N.check(N.lib().v2h_structure_cost_map(self._h, i1, i2))
        return i1, i2

    def gather_stats(self):
        s = np.zeros(8, dtype=np.int64)
        N.check(N.lib().v2h_structure_gather_stats(self._h, s))
        keys = ("a_slices", "a_entries", "a_long_rows", "a_long_entries",
                "at_slices", "at_entries", "at_long_rows", "at_long_entries")
        return dict(zip(keys, s.tolist()))

    def opdm(self, x, type="SUM"):
        """Dense n x n 1-RDM ("A", "B" or "SUM") of primal vector x, caller order."""
        out = np.zeros((self.nmo, self.nmo))
        N.check(N.lib().v2h_structure_opdm(self._h, np.ascontiguousarray(x, dtype=np.float64), type.encode(),
                                           out.reshape(-1)))
        return out

    def tpdm(self, x):
        """Dense spin-free 2-RDM of x, shape (n, n, n, n), [i, k, j, l] = D(ij, kl)."""
        n = self.nmo
        out = np.zeros(n ** 4)
        N.check(N.lib().v2h_structure_tpdm(self._h, np.ascontiguousarray(x, dtype=np.float64), out))
        return out.reshape(n, n, n, n)

    def write_rdm_files(self, x, prefix, active, core=(), which=N.RDM_ALL_FULL):
        """The RDM files of primal vector x (host code; see
        Batch.write_rdm_files).  Returns the record count of each file."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.n_primal:
            raise ValueError("x has the wrong length")
        if len(active) != self.nmo:
            raise ValueError("one full-space label per active orbital needed")
        return N.write_rdm_files(N.lib().v2h_structure_write_rdm_files, (self._h, x), prefix, active, core, which)

    def hf_guess(self):
        """TPDM_GUESS HF primal vector (v2h_structure_hf_guess)."""
        x = np.zeros(self.n_primal)
        N.check(N.lib().v2h_structure_hf_guess(self._h, x))
        return x
