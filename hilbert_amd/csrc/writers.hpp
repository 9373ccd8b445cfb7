// writers.hpp -- hilbert's RDM files for downstream codes: the PSIO records
// of WriteOPDM / WriteTPDM / WriteTPDMSpinFree / WriteActiveTPDM
// (write_opdm.cc:40-119, write_tpdm.cc:40-595) and the IWL TPDM of
// WriteTPDM_IWL (write_tpdm_iwl.cc:45-200), from one instance's primal vector.
//
// Orbital labels are full-space indices: the caller's `active[u]` for active
// orbital u (user order) and `core` for the doubly occupied core orbitals, in
// the order hilbert numbers them (by irrep, then index).  Records follow the
// reference's loop order: active-active by irrep and geminal pairs of
// basis.cc, then core-core, then core-active.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "layout.hpp"
#include "rdm.hpp"

namespace v2h {

struct FullSpace {
    std::vector<int> core;    // full-space labels of the core orbitals
    std::vector<int> active;  // full-space label of each active orbital (user order)
};

// D1a (beta = false) or D1b records, {i, j, value} (write_opdm.cc:40-119).
std::vector<OpdmEntry> psio_opdm(const Layout& L, const double* x, const FullSpace& f, bool beta);
// D2aa (part 0), D2bb (1) or D2ab (2) records, {i, j, k, l, value} =
// D2(ij, kl): with the core contributions (WriteTPDM, write_tpdm.cc:40-260) or
// the active-active part alone (WriteActiveTPDM, write_tpdm.cc:521-595).
std::vector<TpdmEntry> psio_tpdm(const Layout& L, const double* x, const FullSpace& f, int part,
                                 bool active_only = false);
// Spin-free D2 records (WriteTPDMSpinFree, write_tpdm.cc:312-519).
std::vector<TpdmEntry> psio_tpdm_spin_free(const Layout& L, const double* x, const FullSpace& f);
// IWL TPDM values (write_tpdm_iwl.cc:45-200), part 0 aa, 1 bb, 2 ab, labels
// (p, q, r, s) as passed to IWL::write_value: chemists' order, the aa / bb
// values carrying hilbert's factor 1/4.
std::vector<TpdmEntry> iwl_tpdm(const Layout& L, const double* x, const FullSpace& f, int part);

// Which files write_rdm_files writes (hilbert's options that call each writer).
enum RdmFiles : int {
    kRdmTpdmActive = 1,    // TPDM_WRITE: WriteActiveTPDM -> .d2aa .d2bb .d2ab
    kRdmTpdmFull = 2,      // TPDM_WRITE_FULL: WriteTPDM -> .d2aa .d2bb .d2ab
    kRdmTpdmSpinFree = 4,  // TPDM_WRITE_SPIN_FREE: WriteTPDMSpinFree -> .d2
    kRdmOpdmFull = 8,      // OPDM_WRITE_FULL: WriteOPDM -> .d1a .d1b
    kRdmTpdmIwl = 16,      // DERTYPE FIRST: WriteTPDM_IWL -> .aa.iwl .bb.iwl .ab.iwl
    kRdmAll = 31,
};
constexpr int kRdmCounts = 9;

// The files of v2h_batch_write_rdm_files, in the order hilbert writes them
// (v2rdm_solver.cc:1474-1509), each a sequence of keyed PSIO entries:
// <prefix>.d2aa, .d2bb, .d2ab, .d2, .d1a, .d1b, and <prefix>.aa.iwl, .bb.iwl,
// .ab.iwl (the "IWL Buffers" entry).  With both kRdmTpdmActive and
// kRdmTpdmFull the full records replace the active ones, as in hilbert.
// counts[9]: records of D1a, D1b, D2aa, D2bb, D2ab, spin-free D2 and values
// of the three IWL files (after the 1e-14 cutoff); -1 for files not written.
void write_rdm_files(const Layout& L, const double* x, const FullSpace& f, const std::string& prefix, int which,
                     int64_t* counts);

constexpr int kIwlIntsPerBuf = 2980;   // psi4 libiwl buffer length
constexpr double kIwlCutoff = 1.0e-14; // the cutoff hilbert opens its IWL files with

}  // namespace v2h
