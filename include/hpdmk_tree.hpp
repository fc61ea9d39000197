// hpdmk_tree.hpp -- C++ class over the C ABI of libhpdmk_amd.so, with the
// member names and error behaviour of the reference's HPDMKPtTree<Real>
// (include/tree.hpp:27-132): a caller that constructs the reference tree and
// calls form_outgoing_pw / form_incoming_pw / eval_energy / eval_shift_energy /
// update_shift compiles against this class unchanged apart from the
// constructor, which drops the (vestigial, single-rank) sctl::Comm argument.
//
// Real = double keeps the device state in FP64 (HPDMK_FP64); Real = float
// selects mixed precision (FP32 expansions, FP64 phases and sums), which is
// at least as accurate as the reference's float instantiation
// (src/tree.cpp:599-600).  Failures throw std::runtime_error, as the
// reference does (src/tree.cpp:19-28, 353; src/upward_pass.cpp:13).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "hpdmk.h"

namespace hpdmk_amd {

inline void check(int status) {
    if (status != HPDMK_OK) throw std::runtime_error(std::string("hpdmk: ") + hpdmk_last_error());
}

template <typename Real>
class HPDMKPtTree {
  public:
    using float_type = Real;

    // src/tree.cpp:268: builds the octree, kernel tables and device state
    HPDMKPtTree(const HPDMKParams& params, const std::vector<Real>& r_src, const std::vector<Real>& charge)
        : params(params), L(params.L) {
        if (r_src.size() != 3 * charge.size()) throw std::runtime_error("hpdmk: r_src must hold 3 values per charge");
        const std::vector<double> r(r_src.begin(), r_src.end()), q(charge.begin(), charge.end());
        check(hpdmk_tree_create_ex(params, sizeof(Real) == 8 ? HPDMK_FP64 : HPDMK_MIXED, (int)q.size(), r.data(),
                                   q.data(), &h_));
        int info[10];
        check(hpdmk_tree_info(h_, info));
        n_levels_ = info[1];
        n_window = info[2];
        n_diff = info[3];
    }
    HPDMKPtTree(const HPDMKPtTree&) = delete;
    HPDMKPtTree& operator=(const HPDMKPtTree&) = delete;
    ~HPDMKPtTree() {
        if (h_) hpdmk_tree_destroy(h_);
    }

    const HPDMKParams params;
    const Real L;
    int n_window = 0, n_diff = 0;
    int n_levels() const { return n_levels_; }

    // src/planewave.cpp:86-155
    void form_outgoing_pw() { check(hpdmk_form_outgoing_pw(h_)); }
    void form_incoming_pw() { check(hpdmk_form_incoming_pw(h_)); }

    // include/tree.hpp:105-108 and src/energy.cpp:21-173
    Real eval_energy() { return (Real)energy_parts(nullptr); }
    Real eval_energy_window() { return (Real)part(0); }
    Real eval_energy_diff() { return (Real)part(1); }
    Real eval_energy_res() { return (Real)part(2); }

    // src/energy_shift.cpp:18-70; i_unsorted is the caller's original index
    Real eval_shift_energy(long i_unsorted, Real dx, Real dy, Real dz) {
        double dE = 0;
        check(hpdmk_eval_shift_energy(h_, i_unsorted, dx, dy, dz, &dE, nullptr));
        return (Real)dE;
    }

    // src/tree.cpp:447-597; does not require a preceding eval_shift_energy
    void update_shift(long i_unsorted, Real dx, Real dy, Real dz) {
        check(hpdmk_update_shift(h_, i_unsorted, dx, dy, dz));
    }

    // current coordinates in the caller's original order (N x 3)
    std::vector<double> positions(int n_src) {
        std::vector<double> r(3 * (size_t)n_src);
        check(hpdmk_get_positions(h_, r.data()));
        return r;
    }

    // Rebuilds the octree from the current positions (hpdmk_tree_rebuild): the
    // object then holds the rebuilt system; expansions stay formed as before.
    // No counterpart in the reference, which keeps the first topology.
    void rebuild() {
        hpdmk_tree nh = nullptr;
        check(hpdmk_tree_rebuild(h_, &nh));
        hpdmk_tree_destroy(h_);
        h_ = nh;
        int info[10];
        check(hpdmk_tree_info(h_, info));
        n_levels_ = info[1];
    }

    hpdmk_tree handle() const { return h_; }

  private:
    double energy_parts(double* parts) {
        double e = 0;
        check(hpdmk_eval_energy(h_, &e, parts));
        return e;
    }
    double part(int k) {
        double p[3];
        energy_parts(p);
        return p[k];
    }

    hpdmk_tree h_ = nullptr;
    int n_levels_ = 0;
};

}  // namespace hpdmk_amd
